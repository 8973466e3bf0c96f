// Sparse-grid build for one substep on gfx950: hashed block insertion,
// active-block compaction and the particle sort by (block, cell) key.
//
// Replaces WgGrid::launch_sort (src/grid/grid.rs:78-238) and the sort kernels of
// shaders/slosh/grid/{grid,sort}.slang. The active-block set is the reference's
// (primary blocks of every particle plus their 7 "+1" neighbours,
// sort.slang:50-92); header ids are dense but their order differs (the
// reference's order depends on atomic order as well).
//
// Kernel sequence (one launch each, all grid-stride / upper-bound grids so the
// sequence is graph-capturable without reading counts back):
//   k_reset_table      hash slots <- NONE, per-step counters <- 0
//   k_touch_primary    each particle's primary block and its 7 +1 neighbours are
//                      inserted, once per run of equal keys in a wavefront (also
//                      done by the previous substep's fused G2P kernel, which
//                      then replaces this kernel and k_reset_table)
//   k_compact          occupied slots -> dense header ids (workgroup scan, one
//                      atomic per workgroup); zeroes the block's cell counters
//   k_count_nbr        per-particle hash lookup (one per run of equal block keys)
//                      and per-(block, cell) rank via one aggregated atomic per
//                      run of equal (block, cell) keys; in the same launch, per
//                      block the +1 / -1 neighbour and previous-step ids
//   k_block_scan       exclusive scan of the block totals (one workgroup, or
//                      three phases for large capacities; the block totals are
//                      accumulated by k_count_nbr, one atomic per run); the
//                      launch's other workgroups scan each block's 64 cell counts
//   k_scatter          sorted slot = block start + cell offset + rank
#include "common.h"

namespace {

constexpr int ASSOC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                             {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

__global__ __launch_bounds__(256) void k_reset_table(uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                     uint32_t cap, uint32_t *__restrict__ counters) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += stride) {
        keys[s] = SLOSH_NONE;
        vals[s] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[CTR_NUM_BLOCKS] = 0u;
        counters[CTR_SORTED_ALLOC] = 0u;
    }
}

__global__ __launch_bounds__(256) void k_touch_primary(const float4 *__restrict__ pos, uint32_t n, float h,
                                                       uint32_t batched, uint32_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals, uint32_t cap,
                                                       uint32_t *__restrict__ counters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0u, cell;
    if (valid) key = particle_key(pos[i], h, batched, &cell);
    // One insertion round per run of equal keys: the block and its 7 +1 neighbours.
    const uint64_t heads = run_heads(key, valid) & __ballot(valid);
    insert_blocks(keys, cap, batched, key, heads, counters);
}

// Dense header ids for the occupied slots. Each workgroup takes a run of
// 256 * ITEMS slots, ranks its occupied ones with a workgroup scan and reserves
// their ids with one atomic, so the shared counter sees cap / (256 * ITEMS)
// atomics instead of one per wavefront (ITEMS grows with the capacity so the
// number of atomics stays bounded while small tables keep enough workgroups).
template <uint32_t ITEMS>
__global__ __launch_bounds__(256) void k_compact(const uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                 uint32_t cap, uint32_t *__restrict__ hdr_key,
                                                 uint32_t *__restrict__ cell_cnt, uint32_t *__restrict__ blk_cnt,
                                                 uint32_t *__restrict__ counters) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t wg_base;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    constexpr uint32_t SLOTS = 256u * ITEMS;
    for (uint32_t base = blockIdx.x * SLOTS; base < cap; base += gridDim.x * SLOTS) {
        uint32_t k[ITEMS];
        uint32_t mine = 0;
#pragma unroll
        for (uint32_t i = 0; i < ITEMS; i++) {
            const uint32_t s = base + i * 256u + tid;
            k[i] = s < cap ? keys[s] : SLOSH_NONE;
            mine += k[i] != SLOSH_NONE ? 1u : 0u;
        }
        // Workgroup exclusive scan of the per-thread counts.
        uint32_t x = mine;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d);
            if ((int)lane >= d) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; w++) {
            before += w < wave ? wsum[w] : 0u;
            total += wsum[w];
        }
        if (tid == 0) wg_base = total ? atomicAdd(&counters[CTR_NUM_BLOCKS], total) : 0u;
        __syncthreads();
        uint32_t hid = wg_base + before + x - mine;
#pragma unroll
        for (uint32_t i = 0; i < ITEMS; i++) {
            if (k[i] == SLOSH_NONE) continue;
            const uint32_t s = base + i * 256u + tid;
            vals[s] = hid;
            hdr_key[hid] = k[i];
            blk_cnt[hid] = 0u;
            uint4 *cc = reinterpret_cast<uint4 *>(cell_cnt + (size_t)hid * NODES_PER_BLOCK);
#pragma unroll
            for (int j = 0; j < 16; j++) cc[j] = make_uint4(0u, 0u, 0u, 0u);
            hid++;
        }
        __syncthreads();  // wsum / wg_base are reused by the next run
    }
}

__device__ inline void count_particle(uint32_t i, const float4 *__restrict__ pos, uint32_t n, float h, uint32_t batched,
                                               const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                               uint32_t cap, uint32_t *__restrict__ cell_cnt,
                                               uint32_t *__restrict__ pslot, uint32_t *__restrict__ prank,
                                               uint32_t *__restrict__ counters, uint32_t *__restrict__ blk_cnt) {
    const bool valid = i < n;
    const uint32_t lane = lane_id();
    uint32_t cell = 0u, key = 0u;
    bool in_range = true;
    if (valid) {
        const float4 p = pos[i];
        int bx, by, bz;
        assoc_block_cell(h, p.x, p.y, p.z, &bx, &by, &bz, &cell);
        key = pack_key(batched, __float_as_uint(p.w) >> 1, bx, by, bz);
        in_range = block_in_key_range(batched, bx, by, bz);
    }
    // A block outside the key space aliases another one: reported, once per wave.
    if (__ballot(!in_range) && lane == (uint32_t)(__ffsll((unsigned long long)__ballot(!in_range)) - 1))
        atomicOr(&counters[CTR_ERRORS], ERR_KEY_RANGE);
    // Hash lookup by the head of each run of equal block keys, all runs at once.
    const uint64_t bheads = run_heads(key, valid);
    uint32_t hv = SLOSH_NONE;
    if (valid && ((bheads >> lane) & 1ull)) hv = hash_find(keys, vals, cap, key);
    const uint32_t hid = __shfl(hv, run_head_of(bheads));
    const bool ok = valid && hid != SLOSH_NONE;
    // Block totals: one atomic per run of equal block keys.
    if (ok && ((bheads >> lane) & 1ull)) atomicAdd(&blk_cnt[hid], run_length(bheads));
    if (valid && hid == SLOSH_NONE) atomicOr(&counters[CTR_ERRORS], ERR_LOOKUP);
    const uint32_t sk = ok ? hid * NODES_PER_BLOCK + cell : SLOSH_NONE;
    // Per-(block, cell) rank: one atomic per run of equal (block, cell) keys.
    const uint64_t cheads = run_heads(sk, ok);
    uint32_t base = 0u;
    if (ok && ((cheads >> lane) & 1ull)) base = atomicAdd(&cell_cnt[sk], run_length(cheads));
    const int head = run_head_of(cheads);
    base = __shfl(base, head);
    if (valid) {
        pslot[i] = sk;
        prank[i] = base + (lane - (uint32_t)head);
    }
}

// Neighbour table of one block (one wavefront): [0..6] the +1 neighbours in
// the reference's associated-block order (grid.slang:362-382), [7] the block's
// id in the previous substep's table, [8..14] the -1 neighbours.
__device__ inline void block_neighbours(uint32_t hid, uint32_t lane, const uint32_t *__restrict__ hdr_key,
                                        uint32_t *__restrict__ nbr, const uint32_t *__restrict__ keys,
                                        const uint32_t *__restrict__ vals, const uint32_t *__restrict__ prev_keys,
                                        const uint32_t *__restrict__ prev_vals, uint32_t cap, uint32_t batched) {
    if (lane >= NBR_STRIDE) return;
    const uint32_t k = hdr_key[hid];
    int bx, by, bz;
    uint32_t scene;
    unpack_key(batched, k, &bx, &by, &bz, &scene);
    uint32_t out = SLOSH_NONE;
    if (lane < 7) {
        const int *o = ASSOC[lane + 1];
        out = hash_find(keys, vals, cap, pack_key(batched, scene, bx + o[0], by + o[1], bz + o[2]));
    } else if (lane == NBR_PREV) {
        out = hash_find(prev_keys, prev_vals, cap, k);
    } else if (lane >= NBR_MINUS && lane < NBR_MINUS + 7) {
        const int *o = ASSOC[lane - NBR_MINUS + 1];
        out = hash_find(keys, vals, cap, pack_key(batched, scene, bx - o[0], by - o[1], bz - o[2]));
    }
    nbr[(size_t)hid * NBR_STRIDE + lane] = out;
}

// update_block_particle_count (sort.slang:208-246) for the particles, fused with
// update_nbh_block_ids (sort.slang:100-117) for the blocks: both only need the
// header ids, so the first count_wgs workgroups count particles and the rest
// fill neighbour tables (one wavefront per block).
__global__ __launch_bounds__(256) void k_count_nbr(const float4 *__restrict__ pos, uint32_t n, float h,
                                                   uint32_t batched, const uint32_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ vals, uint32_t cap,
                                                   uint32_t *__restrict__ cell_cnt, uint32_t *__restrict__ pslot,
                                                   uint32_t *__restrict__ prank, uint32_t *__restrict__ counters,
                                                   uint32_t *__restrict__ blk_cnt, uint32_t count_wgs,
                                                   const uint32_t *__restrict__ hdr_key,
                                                   uint32_t *__restrict__ nbr, const uint32_t *__restrict__ prev_keys,
                                                   const uint32_t *__restrict__ prev_vals) {
    // The neighbour workgroups come first in dispatch order: each walks a long
    // chain of dependent hash probes, so they start while the counting
    // workgroups (short, independent) fill the rest of the machine.
    const uint32_t nbr_wgs = gridDim.x - count_wgs;
    if (blockIdx.x >= nbr_wgs) {
        count_particle((blockIdx.x - nbr_wgs) * blockDim.x + threadIdx.x, pos, n, h, batched, keys, vals, cap,
                       cell_cnt, pslot, prank, counters, blk_cnt);
        return;
    }
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t stride = nbr_wgs * (blockDim.x >> 6);
    for (uint32_t hid = blockIdx.x * (blockDim.x >> 6) + wave; hid < B; hid += stride)
        block_neighbours(hid, lane, hdr_key, nbr, keys, vals, prev_keys, prev_vals, cap, batched);
}

// Per block (one wavefront): exclusive scan of the 64 cell counts (offsets
// relative to the block's first slot). Waves [w0, w0 + nw) of the grid take
// blocks w0, w0 + nw, ...
__device__ inline void cell_starts(const uint32_t *__restrict__ cell_cnt, uint32_t *__restrict__ cell_start,
                                   uint32_t B, uint32_t w, uint32_t nw) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t hid = w; hid < B; hid += nw) {
        const uint32_t c = cell_cnt[(size_t)hid * NODES_PER_BLOCK + lane];
        uint32_t x = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d);
            if ((int)lane >= d) x += y;
        }
        cell_start[(size_t)hid * NODES_PER_BLOCK + lane] = x - c;
    }
}

// Exclusive scan of blk_cnt[base, base + 4096) by one workgroup of 1024 threads
// (4 consecutive blocks per thread), written to blk_start with `carry` added;
// returns the tile's total. Every thread of the workgroup must call it.
__device__ inline uint32_t scan_block_tile(const uint32_t *__restrict__ blk_cnt, uint32_t *__restrict__ blk_start,
                                           uint32_t base, uint32_t B, uint32_t carry, uint32_t *wsum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t idx = base + tid * 4u;
    const uint32_t v0 = idx < B ? blk_cnt[idx] : 0u;
    const uint32_t v1 = idx + 1 < B ? blk_cnt[idx + 1] : 0u;
    const uint32_t v2 = idx + 2 < B ? blk_cnt[idx + 2] : 0u;
    const uint32_t v3 = idx + 3 < B ? blk_cnt[idx + 3] : 0u;
    const uint32_t t = v0 + v1 + v2 + v3;
    uint32_t x = t;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if ((int)lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            uint32_t y = __shfl_up(w, d);
            if ((int)lane >= d) w += y;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    const uint32_t e = carry + (wave > 0 ? wsum[wave - 1] : 0u) + x - t;
    if (idx < B) blk_start[idx] = e;
    if (idx + 1 < B) blk_start[idx + 1] = e + v0;
    if (idx + 2 < B) blk_start[idx + 2] = e + v0 + v1;
    if (idx + 3 < B) blk_start[idx + 3] = e + v0 + v1 + v2;
    const uint32_t total = wsum[15];
    __syncthreads();
    return total;
}

// Exclusive scan of the block totals in one workgroup: blk_start[b] = sum of
// blk_cnt[< b] (used while the capacity is at most a few tiles).
// Workgroups 1.. compute the per-block cell offsets meanwhile (independent work
// in the same launch).
__global__ __launch_bounds__(1024) void k_block_scan(const uint32_t *__restrict__ blk_cnt,
                                                     uint32_t *__restrict__ blk_start,
                                                     uint32_t *__restrict__ counters,
                                                     const uint32_t *__restrict__ cell_cnt,
                                                     uint32_t *__restrict__ cell_start) {
    __shared__ uint32_t wsum[16];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    if (blockIdx.x > 0) {
        cell_starts(cell_cnt, cell_start, B, (blockIdx.x - 1) * 16u + (threadIdx.x >> 6), (gridDim.x - 1) * 16u);
        return;
    }
    uint32_t carry = 0;
    for (uint32_t base = 0; base < B; base += 4096u) carry += scan_block_tile(blk_cnt, blk_start, base, B, carry, wsum);
    if (threadIdx.x == 0) counters[CTR_SORTED_ALLOC] = carry;
}

// Three-phase scan for large block counts: tiles of 4096 blocks scanned in
// parallel, then the tile totals, then the tile offsets added.
__global__ __launch_bounds__(1024) void k_block_scan_tiles(const uint32_t *__restrict__ blk_cnt,
                                                           uint32_t *__restrict__ blk_start,
                                                           uint32_t *__restrict__ tile_sums,
                                                           const uint32_t *__restrict__ counters,
                                                           uint32_t scan_wgs, const uint32_t *__restrict__ cell_cnt,
                                                           uint32_t *__restrict__ cell_start) {
    __shared__ uint32_t wsum[16];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    if (blockIdx.x >= scan_wgs) {  // the remaining workgroups compute the cell offsets
        cell_starts(cell_cnt, cell_start, B, (blockIdx.x - scan_wgs) * 16u + (threadIdx.x >> 6),
                    (gridDim.x - scan_wgs) * 16u);
        return;
    }
    for (uint32_t t = blockIdx.x; t * 4096u < B; t += scan_wgs) {
        const uint32_t total = scan_block_tile(blk_cnt, blk_start, t * 4096u, B, 0u, wsum);
        if (threadIdx.x == 0) tile_sums[t] = total;
    }
}

__global__ __launch_bounds__(1024) void k_block_scan_top(uint32_t *__restrict__ tile_sums,
                                                         uint32_t *__restrict__ counters) {
    __shared__ uint32_t wsum[16];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    const uint32_t ntiles = (B + 4095u) / 4096u;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 4096u)
        carry += scan_block_tile(tile_sums, tile_sums, base, ntiles, carry, wsum);
    if (threadIdx.x == 0) counters[CTR_SORTED_ALLOC] = carry;
}

__global__ __launch_bounds__(256) void k_block_scan_add(uint32_t *__restrict__ blk_start,
                                                        const uint32_t *__restrict__ tile_sums,
                                                        const uint32_t *__restrict__ counters) {
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    for (uint32_t b = 4096u + blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x)
        blk_start[b] += tile_sums[b >> 12];
}

__global__ __launch_bounds__(256) void k_scatter(const uint32_t *__restrict__ pslot, const uint32_t *__restrict__ prank,
                                                 const uint32_t *__restrict__ cell_start,
                                                 const uint32_t *__restrict__ blk_start, uint32_t *__restrict__ perm,
                                                 uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = pslot[i];
    if (s != SLOSH_NONE) perm[blk_start[s / NODES_PER_BLOCK] + cell_start[s] + prank[i]] = i;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.

struct GridBuffers {
    uint32_t *keys, *vals, *prev_keys, *prev_vals;
    uint32_t *hdr_key, *nbr, *cell_cnt, *cell_start, *blk_start, *blk_cnt;
    uint32_t *pslot, *prank, *perm, *counters;
    uint32_t *tile_sums;  // ceil(cap / 4096) words
};

static inline uint32_t div_up(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }
static inline uint32_t min_u(uint32_t a, uint32_t b) { return a < b ? a : b; }

// `primed`: the previous substep already reset this table and inserted every
// particle's primary block (fused into its P2G and G2P kernels). `ev` (optional)
// brackets the whole sequence.
hipError_t launch_grid_build(const GridBuffers &g, const float4 *pos, uint32_t n, float h, uint32_t batched,
                             uint32_t cap, bool primed, const PassEvents *ev, hipStream_t stream) {
    const uint32_t slot_grid = min_u(div_up(cap, 256), 2048);
    const uint32_t part_grid = div_up(n, 256);
    PassEvents first, last;
    if (ev) {
        first.start = ev->start;
        last.stop = ev->stop;
    }
    const PassEvents *f = ev ? &first : nullptr;
    if (!primed) {
        launch_kernel(f, k_reset_table, dim3(slot_grid), dim3(256), 0, stream, g.keys, g.vals, cap, g.counters);
        f = nullptr;
        if (n > 0)
            launch_kernel(f, k_touch_primary, dim3(part_grid), dim3(256), 0, stream, pos, n, h, batched, g.keys,
                          g.vals, cap, g.counters);
    }
    if (cap <= 65536u)
        launch_kernel(f, k_compact<1>, dim3(min_u(div_up(cap, 256), 2048)), dim3(256), 0, stream,
                      (const uint32_t *)g.keys, g.vals, cap, g.hdr_key, g.cell_cnt, g.blk_cnt, g.counters);
    else
        launch_kernel(f, k_compact<8>, dim3(min_u(div_up(cap, 2048), 2048)), dim3(256), 0, stream,
                      (const uint32_t *)g.keys, g.vals, cap, g.hdr_key, g.cell_cnt, g.blk_cnt, g.counters);
    const uint32_t nbr_wgs = min_u(div_up(cap, 4), 2048);
    launch_kernel(nullptr, k_count_nbr, dim3(part_grid + nbr_wgs), dim3(256), 0, stream, pos, n, h, batched,
                  (const uint32_t *)g.keys, (const uint32_t *)g.vals, cap, g.cell_cnt, g.pslot, g.prank, g.counters,
                  g.blk_cnt, part_grid, (const uint32_t *)g.hdr_key, g.nbr, (const uint32_t *)g.prev_keys,
                  (const uint32_t *)g.prev_vals);
    const PassEvents *scan_ev = n > 0 ? nullptr : (ev ? &last : nullptr);
    const uint32_t cell_wgs = min_u(div_up(cap, 16), 1024);  // 16 blocks (wavefronts) per workgroup
    if (cap <= 16384u) {
        launch_kernel(scan_ev, k_block_scan, dim3(1 + cell_wgs), dim3(1024), 0, stream, (const uint32_t *)g.blk_cnt,
                      g.blk_start, g.counters, (const uint32_t *)g.cell_cnt, g.cell_start);
    } else {
        const uint32_t scan_wgs = min_u(div_up(cap, 4096), 1024);
        launch_kernel(nullptr, k_block_scan_tiles, dim3(scan_wgs + cell_wgs), dim3(1024), 0, stream,
                      (const uint32_t *)g.blk_cnt, g.blk_start, g.tile_sums, (const uint32_t *)g.counters, scan_wgs,
                      (const uint32_t *)g.cell_cnt, g.cell_start);
        launch_kernel(nullptr, k_block_scan_top, dim3(1), dim3(1024), 0, stream, g.tile_sums, g.counters);
        launch_kernel(scan_ev, k_block_scan_add, dim3(min_u(div_up(cap, 256), 4096)), dim3(256), 0, stream,
                      g.blk_start, (const uint32_t *)g.tile_sums, (const uint32_t *)g.counters);
    }
    if (n > 0)
        launch_kernel(ev ? &last : nullptr, k_scatter, dim3(part_grid), dim3(256), 0, stream,
                      (const uint32_t *)g.pslot, (const uint32_t *)g.prank, (const uint32_t *)g.cell_start,
                      (const uint32_t *)g.blk_start, g.perm, n);
    return hipGetLastError();
}
