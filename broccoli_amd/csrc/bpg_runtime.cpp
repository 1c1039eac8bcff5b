// bpg_runtime.cpp -- host runtime shared by the plans (bpg_runtime.h).
#include "bpg_runtime.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace bpg {

float empty_max() {
    const float init = -1000000.0f;
    int m;
    std::memcpy(&m, &init, sizeof m);
    return (float)((float)m / 10000.0f);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

void compact_mask(const float* mask, size_t G, std::vector<unsigned>& comp2grid, std::vector<int>& grid2comp) {
    comp2grid.clear();
    grid2comp.assign(G, -1);
    for (size_t g = 0; g < G; g++)
        if (mask[g] == 1.0f) {
            grid2comp[g] = (int)comp2grid.size();
            comp2grid.push_back((unsigned)g);
        }
}

void build_neighbours(int W, int H, int D, const std::vector<unsigned>& comp2grid,
                      const std::vector<int>& grid2comp, std::vector<int>& nbr) {
    static const int off[BPG_NBR][3] = {
        {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},                                   // faces
        {-1, -1, 0}, {1, -1, 0},                                              // dz = 0 edges
        {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {1, 0, -1},    // dz = -1
        {-1, 1, -1}, {0, 1, -1}, {1, 1, -1}};
    const int Vc = (int)comp2grid.size();
    nbr.assign((size_t)std::max(Vc, 1) * BPG_NBR, -1);
    for (int c = 0; c < Vc; c++) {
        const size_t g = comp2grid[c];
        const int x = (int)(g % W), y = (int)((g / W) % H), z = (int)(g / ((size_t)W * H));
        for (int j = 0; j < BPG_NBR; j++) {
            const int xx = x + off[j][0], yy = y + off[j][1], zz = z + off[j][2];
            if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) continue;
            nbr[(size_t)c * BPG_NBR + j] = grid2comp[xx + (size_t)yy * W + (size_t)zz * W * H];
        }
    }
}

int cluster_max(bpg_ctx* ctx, ClusterScratch& s, const CclArgs& base, long K, unsigned* maxsize, hipStream_t st) {
    const int Vc = base.Vc;
    // union-find scratch of the global path: parent and size per voxel for
    // up to B maps at a time
    const size_t budget = (size_t)8 << 30;
    const size_t per_map = (size_t)std::max(Vc, 1) * 2 * sizeof(unsigned);
    const long B = std::min<long>((long)std::max<size_t>(1, budget / per_map), K);
    BPG_HIP(ctx, hipMemsetAsync(maxsize, 0, K * sizeof(unsigned), st));
    if (K > 0 && ccl_lds_fits(base.nwords, ctx->lds_per_block)) {
        // every map in LDS, two blocks per CU, one launch; the maps with too
        // many active voxels for that (listed by the kernel) take one block
        // per CU, and the few still too large the global path
        CclArgs cc = base;
        cc.maxsize = maxsize;
        cc.labels = nullptr;
        cc.map_ids = nullptr;
        cc.nmaps = (int)K;
        BPG_HIP(ctx, s.overflow.ensure((size_t)2 * K + 2));
        unsigned* list1 = s.overflow.p;          // [K] maps left by the first tier
        unsigned* list2 = s.overflow.p + K;      // [K] maps left by the second
        unsigned* counts = s.overflow.p + 2 * K; // [2]
        BPG_HIP(ctx, hipMemsetAsync(counts, 0, 2 * sizeof(unsigned), st));
        BPG_HIP(ctx, launch_ccl_lds(cc, ctx->lds_per_block, 2, nullptr, (int)K, list1, counts, st));
        unsigned n_over[2] = {0, 0};
        BPG_HIP(ctx, hipMemcpyAsync(n_over, counts, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        BPG_HIP(ctx, hipStreamSynchronize(st));
        if (n_over[0] > 0) {
            BPG_HIP(ctx, launch_ccl_lds(cc, ctx->lds_per_block, 1, list1, (int)n_over[0], list2, counts + 1, st));
            BPG_HIP(ctx, hipMemcpyAsync(n_over + 1, counts + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
            BPG_HIP(ctx, hipStreamSynchronize(st));
        }
        if (n_over[1] > 0) {
            const long B2 = std::min<long>(B, n_over[1]);
            BPG_HIP(ctx, s.parent.ensure((size_t)B2 * std::max(Vc, 1)));
            BPG_HIP(ctx, s.sizes.ensure((size_t)B2 * std::max(Vc, 1)));
            cc.parent = s.parent.p;
            cc.sizes = s.sizes.p;
            for (long j0 = 0; j0 < (long)n_over[1]; j0 += B2) {
                cc.map_ids = list2 + j0;
                cc.nmaps = (int)std::min<long>(B2, (long)n_over[1] - j0);
                BPG_HIP(ctx, launch_ccl(cc, st));
            }
        }
        return BPG_OK;
    }
    BPG_HIP(ctx, s.parent.ensure((size_t)B * std::max(Vc, 1)));
    BPG_HIP(ctx, s.sizes.ensure((size_t)B * std::max(Vc, 1)));
    for (long k0 = 0; k0 < K; k0 += B) {
        CclArgs cc = base;
        cc.bits = base.bits + (size_t)k0 * base.nwords;
        cc.parent = s.parent.p;
        cc.sizes = s.sizes.p;
        cc.maxsize = maxsize + k0;
        cc.labels = nullptr;
        cc.values = base.values ? base.values + (size_t)k0 * Vc : nullptr;
        cc.map_ids = nullptr;
        cc.nmaps = (int)std::min<long>(B, K - k0);
        BPG_HIP(ctx, launch_ccl(cc, st));
    }
    return BPG_OK;
}

}  // namespace bpg
