// Host-side launch interfaces shared by the kernel translation units
// (kernels_grid.hip, kernels_transfer.hip, kernels_particles.hip) and the C ABI
// implementation (slosh_api.hip): the device buffers each launcher takes and
// the launchers themselves.
#pragma once

#include "common.h"

// ---- grid sort (kernels_grid.hip) ---------------------------------------------
struct GridBuffers {
    uint32_t *keys, *prev_keys;
    uint32_t *blist, *nbr, *cell_cnt, *cell_start, *blk_start, *blk_cnt;
    uint32_t *pslot, *prank, *perm, *counters;
    uint32_t *tile_sums;  // ceil(cap / 4096) words
    // Movers listed by the previous substep and their per-block lists.
    const uint32_t *mv_key, *mv_slot;
    uint32_t *mcount, *mlist, *ovf_slot, *ovf_hid;
    // Previous table's block ranges and stayer counts, and per block of this
    // table those of its previous-table block (written by the steady-state sort).
    const uint32_t *prev_start, *prev_cnt, *prev_stay;
    uint4 *bin;
};

// `fast`: the previous substep's fused G2P kernel filled this table, wrote every
// block's stayers first and listed the movers. `ev` (optional) brackets the
// whole sequence.
hipError_t launch_grid_build(const GridBuffers &g, const float4 *pos, uint32_t n, float h, uint32_t batched,
                             uint32_t cap, bool fast, const PassEvents *ev, hipStream_t stream);

// ---- transfers (kernels_transfer.hip) ------------------------------------------
struct TransferBuffers {
    uint32_t fast;  // steady-state substep (members from stayer ranges and mover lists)
    const uint32_t *keys, *blist, *nbr, *counters;
    const uint32_t *perm;                             // full sort
    uint32_t *cell_cnt, *blk_start, *blk_cnt, *stay;  // this table
    const uint4 *bin;  // per block: range (start, count) and stayers of its previous-table block
    const uint32_t *mcount, *mlist, *ovf_slot, *ovf_hid;
    uint32_t *errors;  // this substep's error word
    float4 *tiles, *nodes;
    void *coll_prev, *coll_cur;
    float *fdt;
};

struct ParticleBuffers {
    float4 *pos, *vm;
    float *C, *F, *vgd, *dps;
    uint32_t *ids;
};

struct StaticBuffers {
    const float4 *prop, *model, *dp0, *dp1;
    const float4 *scene_rec;  // batched: 4 records (prop, model, dp0, dp1) per scene id
};

struct ColliderBuffers {
    const float4 *shapes, *poses;
    const uint2 *mats;
    uint32_t count;
    const float4 *mesh_vtx;    // trimesh collision vertices (x, y, z, pad)
    const uint32_t *mesh_idx;  // trimesh collision indices
};

struct BodyBuffers {
    float4 *poses, *vels;
    const float4 *mprops, *local_com;
    int *impulses;
    uint32_t count;
};

// The next substep's table and sort inputs, prepared by this substep's P2G
// (reset) and fused G2P (insertion, mover list).
struct NextTableBuffers {
    uint32_t *keys, *counters;
    uint32_t *mcount;            // next table: movers per block
    uint32_t *mv_key, *mv_slot;  // movers: new block key, output slot
    uint32_t prime;
};

hipError_t launch_p2g(const StepParams &P, const ParticleBuffers &pb, uint32_t pcap, const TransferBuffers &t,
                      const NextTableBuffers &next, const PassEvents *ev, hipStream_t stream);
hipError_t launch_grid_update(const StepParams &P, const TransferBuffers &t, const ColliderBuffers &c, uint32_t mode,
                              uint32_t do_update, const PassEvents *ev, hipStream_t stream);
hipError_t launch_g2p_update(const StepParams &P, const ParticleBuffers &in, const ParticleBuffers &out,
                             const StaticBuffers &s, uint32_t pcap, const TransferBuffers &t,
                             const NextTableBuffers &next, const BodyBuffers &bb, const PassEvents *ev,
                             hipStream_t stream);
hipError_t launch_g2p(const StepParams &P, const ParticleBuffers &b, uint32_t pcap, const TransferBuffers &t,
                      const PassEvents *ev, hipStream_t stream);
hipError_t launch_particle_update(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s, uint32_t pcap,
                                  float *fdt, const PassEvents *ev, hipStream_t stream);
hipError_t launch_rigid_update(const StepParams &P, const BodyBuffers &bb, const PassEvents *ev, hipStream_t stream);
// Per-scene record table: for each run of equal scene ids among storage slots
// [slot0, slot0 + n) of a particle buffer, the by-id records of the run's first
// particle (read only while every scene has one record), then each scene's mass
// (the largest of those particles' masses) in the plastic record's z (read only
// while every scene's particles have mass = enabled ? its mass : +0).
hipError_t launch_scene_records(const float4 *pos, const float4 *vm, const uint32_t *ids, uint32_t slot0, uint32_t n,
                                const StaticBuffers &s, float4 *scene_rec, hipStream_t stream);
hipError_t launch_timestep_bound(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s, uint32_t pcap,
                                 uint32_t *out, hipStream_t stream);
// Two passes over partial minima, one per TSB_CHUNK particles, in `partials`
// (div_up(n, TSB_CHUNK) words); *out = the bound, 0xffffffff with no enabled
// particle. Neither buffer needs clearing first.
hipError_t launch_timestep_bound_partials(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s,
                                          uint32_t pcap, uint32_t *partials, uint32_t *out, hipStream_t stream);

// ---- particle storage (kernels_particles.hip) ------------------------------------
struct CompactArgs {
    // current and destination particle sets (slot order)
    float4 *pos[2], *vm[2];
    float *C[2], *F[2], *vgd[2], *dps[2], *fdt[2];
    uint32_t *ids[2];
    // by-id static records, current and destination
    float4 *prop[2], *model[2], *dp0[2], *dp1[2];
    uint32_t *keep_slot, *keep_id, *keep_id2, *tile_sums, *total;
};
hipError_t launch_compact(const CompactArgs &a, uint32_t n, uint32_t pcap, uint32_t mode, const uint8_t *dead,
                          uint32_t lo, uint32_t hi, hipStream_t st);

struct GatherArgs {
    float4 *pos, *vm;
    float *C, *F, *vgd, *dps, *fdt;
    uint32_t *ids;
    // staging (count entries per component)
    float4 *s_pos, *s_vm;
    float *s_C, *s_F, *s_vgd, *s_dps, *s_fdt;
    uint32_t *rank_id, *tile_sums, *total;
};
hipError_t launch_scene_gather(const GatherArgs &a, uint32_t n, uint32_t pcap, uint32_t scene, uint32_t count,
                               hipStream_t st);
uint32_t scan_tiles(uint32_t n);
