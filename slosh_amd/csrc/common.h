// Shared device definitions of the gfx950 MPM substep.
//
// Grid vocabulary follows the reference (shaders/slosh/grid/grid.slang):
// blocks of 4x4x4 nodes and a virtual block id per block. A hash map from the
// packed block key gives each active block a "header id" (hid): here the slot
// the key occupies in the substep's table, so a lookup needs no value array
// and no compaction pass assigns ids. Per-block arrays are indexed by slot; the
// block list (slots of the occupied entries, one per block) gives the order the
// per-block kernels walk and the order of particle storage.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#define SLOSH_NONE 0xffffffffu

// Number of nodes per block (grid.slang:31) and of nodes in a block's
// quadratic-stencil footprint (the block plus the +1 halo: 6 x 6 x 6).
#define NODES_PER_BLOCK 64
#define TILE_NODES 216

// A block's P2G tile (its 4^3 nodes and the +1 halo, 6^3 nodes) is stored region
// by region, so that what the grid update reads from each neighbour's tile is
// contiguous: region r = (x >= 4) | (y >= 4) << 1 | (z >= 4) << 2 of extent 2
// along the axes where it lies in the halo and 4 along the others.
__device__ __host__ inline uint32_t tile_region_base(uint32_t r) {
    // sizes 64, 32, 32, 16, 32, 16, 16, 8
    return r == 0u ? 0u : r == 1u ? 64u : r == 2u ? 96u : r == 3u ? 128u : r == 4u ? 144u : r == 5u ? 176u
                                                                                             : r == 6u ? 192u : 208u;
}

__device__ __host__ inline uint32_t tile_index(uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t ox = x >= 4u, oy = y >= 4u, oz = z >= 4u;
    const uint32_t ex = ox ? 2u : 4u, ey = oy ? 2u : 4u;
    return tile_region_base(ox | (oy << 1) | (oz << 2)) + (x - 4u * ox) + ex * ((y - 4u * oy) + ey * (z - 4u * oz));
}

// Counter slots of one substep's counter set (two sets alternate with the hash
// tables so a substep can prepare the next one's).
#define CTR_NUM_BLOCKS 0
#define CTR_SORTED_ALLOC 1
#define CTR_ERRORS 2
#define CTR_MOVERS 3     // entries of the mover list destined for this table
#define CTR_OVERFLOW 4   // movers beyond a block's MOVER_CAP list slots
#define ERR_HASH_FULL 1u
#define ERR_LOOKUP 2u
#define ERR_KEY_RANGE 4u

// Particle storage between substeps: every block's particles are contiguous. A
// particle that leaves its block (a "mover") is listed with its new block key by
// the substep that moved it; the next substep's sort places the movers of each
// block in MOVER_CAP list slots per block (and any beyond them in a shared
// overflow list).
#define MOVER_CAP 64

// Particles per workgroup of the two-pass timestep bound (one partial minimum each).
#define TSB_CHUNK 2048u

// Neighbour table per block (16 words): [0..6] the +1 neighbours in the
// reference's associated-block order (grid.slang:371-379, offsets 1..7),
// [7] the block's header id in the previous substep's hash map,
// [8..14] the -1 neighbours (same offsets, negated), [15] unused.
#define NBR_STRIDE 16
#define NBR_PREV 7
#define NBR_MINUS 8

struct StepParams {
    float gravity[3];
    float dt;
    float h;      // cell width
    float inv_h;  // 1.0f / h, as p2g_scatter_style.slang:54
    float h_dt;   // h / dt (FP32 division, as the velocity clamp computes it)
    float inv_d;  // 4.0f / (h * h), the APIC inverse inertia factor
    uint32_t cap; // hash capacity (power of two) == max blocks
    uint32_t n;   // particle count
    uint32_t batched;     // 0: pack_key of grid.slang:83-90; 1: scene << 18 | 6-bit x, y, z (offset 16)
    uint32_t num_bodies;
    uint32_t has_force;   // force_dt may be non-zero
    uint32_t has_sand;    // DP state arrays allocated
    uint32_t coll_cache;  // use the per-node collision cache (moving colliders)
    uint32_t cur;         // ping-pong index of the current hash table / cache
    uint32_t svd_f64;     // SVD-based stress / plasticity in FP64 (1) or FP32 (0)
    uint32_t tag_mask;    // material tags present (bit t = tag t; bit 4 = any other tag)
    uint32_t uniform;     // every particle has the same static record: use u_* below
    float4 u_prop, u_model, u_dp0, u_dp1;
    // Batched data whose every scene has one static record: read it from the
    // per-scene table (scene id from the position's w bits) instead of by id.
    uint32_t scene_uniform;
    // Batched data whose every scene's particles were given with mass = enabled ?
    // m_scene : +0 (and not exposed since): the mass comes from the scene table.
    uint32_t scene_mass;
    // Every particle was created enabled with mass u_mass, and only the update has
    // disabled particles since (it zeroes their mass): mass = enabled ? u_mass : 0.
    uint32_t mass_uniform;
    float u_mass;
    // Launch shape (autotuned): particles staged in LDS per P2G round, threads per
    // G2P/update workgroup, workgroups of the per-block kernels.
    uint32_t p2g_chunk, g2p_threads, block_wgs;
};

// ---------------------------------------------------------------------------
// Block keys.

// pack_key, grid.slang:83-90
__device__ __host__ inline uint32_t pack_key_single(int x, int y, int z) {
    return ((uint32_t)(x + 0x3ff) & 0x7ffu) | (((uint32_t)(y + 0x1ff) & 0x3ffu) << 11) |
           (((uint32_t)(z + 0x3ff) & 0x7ffu) << 21);
}

__device__ __host__ inline uint32_t pack_key_batched(uint32_t scene, int x, int y, int z) {
    return (scene << 18) | (((uint32_t)(z + 16) & 63u) << 12) | (((uint32_t)(y + 16) & 63u) << 6) |
           ((uint32_t)(x + 16) & 63u);
}

// Block coordinates whose key and whose "+1" neighbours' keys are exact: batched
// keys hold 6 bits per axis with offset 16 ([-16, 47]), single-scene keys 11 / 10
// / 11 bits with offsets 0x3ff / 0x1ff / 0x3ff (grid.slang:83-90). A block outside
// this range would alias another block of the same key space.
__device__ __host__ inline bool block_in_key_range(uint32_t batched, int x, int y, int z) {
    if (batched) return x >= -16 && x <= 46 && y >= -16 && y <= 46 && z >= -16 && z <= 46;
    return x >= -1023 && x <= 1023 && y >= -511 && y <= 511 && z >= -1023 && z <= 1023;
}

__device__ __host__ inline uint32_t pack_key(uint32_t batched, uint32_t scene, int x, int y, int z) {
    return batched ? pack_key_batched(scene, x, y, z) : pack_key_single(x, y, z);
}

__device__ __host__ inline void unpack_key(uint32_t batched, uint32_t k, int *x, int *y, int *z, uint32_t *scene) {
    if (batched) {
        *x = (int)(k & 63u) - 16;
        *y = (int)((k >> 6) & 63u) - 16;
        *z = (int)((k >> 12) & 63u) - 16;
        *scene = k >> 18;
    } else {
        *x = (int)(k & 0x7ffu) - 0x3ff;
        *y = (int)((k >> 11) & 0x3ffu) - 0x1ff;
        *z = (int)((k >> 21) & 0x7ffu) - 0x3ff;
        *scene = 0;
    }
}

// Slot hash: the Murmur3 key mixing of grid.slang:93-100 followed by Murmur3's
// 32-bit finalizer (fmix32). Without the finalizer the low bits of the mix keep
// the lattice structure of block keys: on the 1M-particle dam break (3400
// blocks in 16384 slots) linear probing then needs 3.5 probes per lookup on
// average and up to 22, against 1.2 and at most 7 with it. Which slots the keys
// occupy changes only the table's internal order (header ids); the block set,
// the per-block counts and every particle's result do not depend on it.
__device__ __host__ inline uint32_t hash_key(uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    k ^= k >> 16;
    k *= 0x85ebca6bu;
    k ^= k >> 13;
    k *= 0xc2b2ae35u;
    k ^= k >> 16;
    return k;
}

// Associated cell (round(x / h) - 1, half-to-even as WGSL round) and the block
// it belongs to: block_associated_to_point (grid.slang:346-355) and
// associated_cell_index_in_block_off_by_one (particle.slang:79-87).
// x / h. For a power-of-two h (normal, with a normal reciprocal) the product
// with the exact reciprocal 2^-e is the same number as the quotient: both are
// exact scalings, or round the same exact value. The test is uniform (h is a
// kernel argument), so the other path costs one scalar branch.
__device__ inline float div_h(float x, float h) {
    const uint32_t hb = __float_as_uint(h);
    if ((hb & 0x807fffffu) == 0u && hb - 0x00800000u < 0x7e000000u)
        return x * __uint_as_float(0x7f000000u - (hb & 0x7f800000u));
    return x / h;
}

__device__ inline void assoc_block_cell(float h, float px, float py, float pz, int *bx, int *by, int *bz,
                                        uint32_t *cell) {
    float ax = rintf(div_h(px, h)) - 1.0f;
    float ay = rintf(div_h(py, h)) - 1.0f;
    float az = rintf(div_h(pz, h)) - 1.0f;
    float fx = floorf(ax / 4.0f), fy = floorf(ay / 4.0f), fz = floorf(az / 4.0f);
    *bx = (int)fx;
    *by = (int)fy;
    *bz = (int)fz;
    uint32_t lx = (uint32_t)(ax - fx * 4.0f);
    uint32_t ly = (uint32_t)(ay - fy * 4.0f);
    uint32_t lz = (uint32_t)(az - fz * 4.0f);
    *cell = lx + 4u * ly + 16u * lz;
}

// Bounded open-addressing lookup (find_block_header_id, grid.slang:176-198);
// returns the key's slot, which is the block's header id.
__device__ inline uint32_t hash_find(const uint32_t *__restrict__ keys, uint32_t cap, uint32_t key) {
    uint32_t slot = hash_key(key) & (cap - 1u);
    for (uint32_t k = 0; k < cap; k++) {
        uint32_t st = keys[slot];
        if (st == key) return slot;
        if (st == SLOSH_NONE) return SLOSH_NONE;
        slot = (slot + 1u) & (cap - 1u);
    }
    return SLOSH_NONE;
}

// Optional per-pass timing: the dispatch packets of the pass's first / last
// kernel write the start / stop timestamps (hipExtLaunchKernelGGL), so the
// measured time is the kernels' own even when the host enqueues slower than the
// GPU executes.
struct PassEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

template <typename... KArgs, typename... Args>
inline void launch_kernel(const PassEvents *ev, void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shmem,
                          hipStream_t stream, Args... args) {
    if (ev && (ev->start || ev->stop))
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, ev->start, ev->stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
}

// Wave-level helpers (64-lane wavefronts).
__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline uint64_t lanemask_lt() { return __lanemask_lt(); }

// Per-block counters a new block must start from zero, cleared by the thread
// whose insertion created the block (null pointers are skipped): its mover
// count, and for the full sort its particle count and 64 cell counts.
struct BlockInit {
    uint32_t *mcount, *blk_cnt, *cell_cnt;
};

// Open-addressing insertion with a strong CAS (insertion_index, grid.slang:118-174).
// A plain load first: most insertions find their key already present (many
// runs of a wavefront, and many wavefronts, touch the same block), which then
// costs no atomic. A stale NONE only leads to the CAS, which sees the truth.
__device__ inline void hash_insert(uint32_t *__restrict__ keys, uint32_t cap, uint32_t key,
                                   uint32_t *__restrict__ counters, const BlockInit &init) {
    uint32_t slot = hash_key(key) & (cap - 1u);
    for (uint32_t k = 0; k < cap; k++) {
        const uint32_t cur = keys[slot];
        if (cur == key) return;
        if (cur == SLOSH_NONE) {
            const uint32_t old = atomicCAS(&keys[slot], SLOSH_NONE, key);
            if (old == SLOSH_NONE) {
                if (init.mcount) init.mcount[slot] = 0u;
                if (init.blk_cnt) init.blk_cnt[slot] = 0u;
                if (init.cell_cnt) {
                    uint4 *cc = reinterpret_cast<uint4 *>(init.cell_cnt + (size_t)slot * NODES_PER_BLOCK);
#pragma unroll
                    for (int j = 0; j < 16; j++) cc[j] = make_uint4(0u, 0u, 0u, 0u);
                }
                return;
            }
            if (old == key) return;
        }
        slot = (slot + 1u) & (cap - 1u);
    }
    atomicOr(&counters[CTR_ERRORS], ERR_HASH_FULL);
}

// Associated blocks of a block (grid.slang:362-382): offset 0 is the block
// itself, 1..7 its "+1" neighbours.
__device__ constexpr int BLOCK_ASSOC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                                              {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

// touch_primary_blocks + touch_neighbor_blocks (sort.slang:50-92) for the keys
// of the lanes in `heads` (a wavefront-uniform lane mask): every head's block
// and its 7 "+1" neighbours are inserted, with the 64 lanes spread over
// (head, offset) pairs, eight heads per round. Every lane of the wavefront must
// call it.
__device__ inline void insert_blocks(uint32_t *__restrict__ keys, uint32_t cap, uint32_t batched, uint32_t key,
                                     uint64_t heads, uint32_t *__restrict__ counters, const BlockInit &init) {
    const uint32_t lane = __lane_id();
    const uint32_t r = lane >> 3, o = lane & 7u;
    while (heads) {
        uint64_t m = heads;
        for (uint32_t k = 0; k < r && m; k++) m &= m - 1ull;
        const int src = m ? __ffsll((unsigned long long)m) - 1 : 0;
        const uint32_t hk = __shfl(key, src);
        if (m) {
            int x, y, z;
            uint32_t scene;
            unpack_key(batched, hk, &x, &y, &z, &scene);
            hash_insert(keys, cap,
                        pack_key(batched, scene, x + BLOCK_ASSOC[o][0], y + BLOCK_ASSOC[o][1], z + BLOCK_ASSOC[o][2]),
                        counters, init);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) heads &= heads ? heads - 1ull : 0ull;
    }
}

// Primary-block key of a particle (x, y, z, w = scene << 1 | enabled bits).
__device__ inline uint32_t particle_key(const float4 &p, float h, uint32_t batched, uint32_t *cell) {
    int bx, by, bz;
    assoc_block_cell(h, p.x, p.y, p.z, &bx, &by, &bz, cell);
    uint32_t scene = __float_as_uint(p.w) >> 1;
    return pack_key(batched, scene, bx, by, bz);
}

// Runs of equal keys over consecutive lanes of a wavefront (particle storage is
// sorted by block and cell, so most wavefronts hold a few long runs). Returns
// the mask of run heads; invalid lanes are heads of their own (ignored) runs.
// Every lane of the wavefront must call it.
__device__ inline uint64_t run_heads(uint32_t key, bool valid) {
    const uint32_t lane = lane_id();
    const uint32_t prev = __shfl_up(key, 1);
    const int prev_valid = __shfl_up((int)valid, 1);
    const bool head = !valid || lane == 0 || !prev_valid || prev != key;
    return __ballot(head);
}

// Lane index of the head of this lane's run.
__device__ inline int run_head_of(uint64_t heads) {
    const uint64_t le = heads & (lanemask_lt() | (1ull << lane_id()));
    return 63 - __clzll((long long)le);
}

// Length of the run starting at this (head) lane.
__device__ inline uint32_t run_length(uint64_t heads) {
    const uint32_t lane = lane_id();
    const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
    const uint32_t next = above ? (uint32_t)(__ffsll((unsigned long long)above) - 1) : 64u;
    return next - lane;
}

