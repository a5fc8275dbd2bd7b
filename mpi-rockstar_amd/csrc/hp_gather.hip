// hp_gather.hip -- record assembly for a batch of calls: the reference's
// calc_particle_radii / _reset_potentials (src/groupies.c:212-268), which
// fills each halo's struct potential records from the FOF group's particle
// copies before compute_potential.
//
// The host walks the halo hierarchy exactly as the reference recursion does
// (the halo, its descendants depth first in sibling order, then the parent's
// subtree as potential-only records where the host rule holds) and emits one
// segment per visited halo: a contiguous run of particles that becomes a
// contiguous run of records.  The device side is one gather over tiles of
// 1024 records: read a 32-byte particle, write a 48-byte record with r2
// about the call's centre.  HBM-bound.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

struct GParticle {  // struct particle, src/particle.h:5-8
    int64_t id;
    float pos[6];
};
static_assert(sizeof(GParticle) == 32, "particle layout");

struct GatherSeg {
    int64_t src;     // first particle
    int64_t dst;     // first record
    int32_t n;       // particles
    int32_t call;    // the call whose centre r2 is about
    int32_t flags;   // 0, 1 (the call's own halo) or 2 (a small descendant)
    int32_t potential_only;
};

struct GatherTile {
    int32_t seg;
    int32_t off;
};

constexpr int kGatherTile = 1024;

// src/groupies.c:212-233 per record: zeroed record, position and velocity,
// r2 = sum of float dx * dx about the centre, ke = -1 for potential-only
// records, flags.
__global__ __launch_bounds__(256) void k_gather_records(const GParticle *__restrict__ parts,
                                                        const GatherSeg *__restrict__ segs,
                                                        const GatherTile *__restrict__ tiles,
                                                        int32_t ntiles,
                                                        const float *__restrict__ cen,
                                                        struct potential *__restrict__ po) {
    for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const GatherTile tl = tiles[t];
        const GatherSeg sg = segs[tl.seg];
        const float c[3] = {cen[3 * sg.call], cen[3 * sg.call + 1], cen[3 * sg.call + 2]};
        const int32_t end = min(tl.off + kGatherTile, sg.n);
        for (int32_t i = tl.off + (int32_t)threadIdx.x; i < end; i += blockDim.x) {
            const GParticle p = parts[sg.src + i];
            struct potential r;
            float r2 = 0.0f;
            for (int k = 0; k < 3; k++) {
                const float dx = p.pos[k] - c[k];
                r2 += dx * dx;
            }
            for (int k = 0; k < 6; k++) r.pos[k] = p.pos[k];
            r.r2 = r2;
            r.pe = 0.0;
            r.ke = sg.potential_only ? -1.0f : 0.0f;
            r.flags = sg.flags;
            po[sg.dst + i] = r;
        }
    }
}

}  // namespace hp
