// Sparse-grid build for one substep on gfx950: hashed block insertion, the
// block list and neighbour tables, and the particle sort by block.
//
// Replaces WgGrid::launch_sort (src/grid/grid.rs:78-238) and the sort kernels of
// shaders/slosh/grid/{grid,sort}.slang. The active-block set is the reference's
// (primary blocks of every particle plus their 7 "+1" neighbours,
// sort.slang:50-92). A block's header id is the slot of its key in the
// substep's hash table; the order of the block list differs from the
// reference's header order (which depends on atomic order as well).
//
// Steady state (the previous substep's fused G2P kernel inserted every
// particle's primary block and its +1 neighbours into this substep's table, wrote
// each block's particles contiguously with the ones that stay in the block first,
// and listed the particles that left their block with their new block key):
//   k_sort_fast        one launch: the block list (occupied slots, workgroup
//                      scan and one atomic per workgroup), each block's +1 / -1
//                      neighbour ids and its id in the previous table, and the
//                      movers bucketed into per-block lists (one lookup and one
//                      atomic per run of equal keys). A block's particles are
//                      then the stayers of its previous-substep range plus its
//                      movers; the per-block kernels read them from there.
// Full sort (first substep, after host mutations, with hooks):
//   k_reset_table      hash slots <- NONE, per-step counters <- 0
//   k_touch_primary    each particle's primary block and its 7 +1 neighbours are
//                      inserted, once per run of equal keys in a wavefront; the
//                      thread that creates a block clears its counters
//   k_sort_fast        (block list and neighbour tables only)
//   k_count            per-particle hash lookup (one per run of equal block keys)
//                      and per-(block, cell) rank via one aggregated atomic per
//                      run of equal (block, cell) keys
//   k_block_scan       exclusive scan of the block totals in block-list order
//                      (one workgroup, or three phases for large capacities); the
//                      launch's other workgroups scan each block's 64 cell counts
//   k_scatter          sorted slot = block start + cell offset + rank
#include "launch.h"

namespace {

constexpr int ASSOC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                             {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

__global__ __launch_bounds__(256) void k_reset_table(uint32_t *__restrict__ keys, uint32_t cap,
                                                     uint32_t *__restrict__ counters) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += stride) keys[s] = SLOSH_NONE;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[CTR_NUM_BLOCKS] = 0u;
        counters[CTR_SORTED_ALLOC] = 0u;
        counters[CTR_MOVERS] = 0u;
        counters[CTR_OVERFLOW] = 0u;
    }
}

__global__ __launch_bounds__(256) void k_touch_primary(const float4 *__restrict__ pos, uint32_t n, float h,
                                                       uint32_t batched, uint32_t *__restrict__ keys, uint32_t cap,
                                                       uint32_t *__restrict__ counters, BlockInit init) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0u, cell;
    if (valid) key = particle_key(pos[i], h, batched, &cell);
    // One insertion round per run of equal keys: the block and its 7 +1 neighbours.
    const uint64_t heads = run_heads(key, valid) & __ballot(valid);
    insert_blocks(keys, cap, batched, key, heads, counters, init);
}

struct MoverLists {
    const uint32_t *key, *slot;  // movers listed by the previous substep (new block key, storage slot)
    uint32_t *mcount, *mlist;    // per block: movers received, the first MOVER_CAP of them
    uint32_t *ovf_slot, *ovf_hid;  // the others, with their block
    // Steady state: per block of the previous table its storage range and stayer
    // count; per block of this table the same three of its previous-table block
    // (null in the full sort).
    const uint32_t *prev_start, *prev_cnt, *prev_stay;
    uint4 *bin;
};

// Neighbour table entry `l` (0..15) of block `hid`: [0..6] the +1 neighbours in
// the reference's associated-block order (grid.slang:362-382), [7] the block's
// id in the previous substep's table, [8..14] the -1 neighbours. With the
// previous table's block ranges given, lane 7 also records the block's range in
// the previous substep's storage and its stayer count.
__device__ inline void block_neighbour(uint32_t hid, uint32_t key, uint32_t l, uint32_t *__restrict__ nbr,
                                       const uint32_t *__restrict__ keys, const uint32_t *__restrict__ prev_keys,
                                       uint32_t cap, uint32_t batched, const MoverLists &mv) {
    int bx, by, bz;
    uint32_t scene;
    unpack_key(batched, key, &bx, &by, &bz, &scene);
    uint32_t out = SLOSH_NONE;
    if (l < 7) {
        const int *o = ASSOC[l + 1];
        out = hash_find(keys, cap, pack_key(batched, scene, bx + o[0], by + o[1], bz + o[2]));
    } else if (l == NBR_PREV) {
        out = hash_find(prev_keys, cap, key);
        if (mv.bin) {
            uint4 r = make_uint4(0u, 0u, 0u, 0u);
            if (out != SLOSH_NONE) r = make_uint4(mv.prev_start[out], mv.prev_cnt[out], mv.prev_stay[out], 0u);
            mv.bin[hid] = r;
        }
    } else if (l >= NBR_MINUS && l < NBR_MINUS + 7) {
        const int *o = ASSOC[l - NBR_MINUS + 1];
        out = hash_find(keys, cap, pack_key(batched, scene, bx - o[0], by - o[1], bz - o[2]));
    }
    nbr[(size_t)hid * NBR_STRIDE + l] = out;
}


// The sort of a steady-state substep (update_block_particle_count and
// update_nbh_block_ids, sort.slang:100-246, for a table the previous substep
// filled), one launch of three kinds of workgroups, in dispatch order:
// [0, mv_wgs) the listed movers, each appended to its new block's list (one
// lookup and one atomic per run of equal keys; those beyond a block's MOVER_CAP
// list slots go to the overflow list with their block); [mv_wgs, mv_wgs +
// cmp_wgs) the block list: each takes 256 * ITEMS slots and appends the occupied
// ones (workgroup scan, one atomic); the rest the neighbour tables, 16 lanes per
// slot, every occupied slot's 15 lookups at once.
template <uint32_t ITEMS>
__global__ __launch_bounds__(256) void k_sort_fast(const uint32_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ prev_keys, uint32_t cap,
                                                   uint32_t batched, uint32_t *__restrict__ counters,
                                                   uint32_t *__restrict__ blist, uint32_t *__restrict__ nbr,
                                                   uint32_t mv_wgs, uint32_t cmp_wgs, MoverLists mv) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t wg_base;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    if (blockIdx.x < mv_wgs) {
        // Movers. The loop bound is wavefront-uniform (run detection uses every lane).
        const uint32_t nmov = counters[CTR_MOVERS];
        const uint32_t stride = mv_wgs * 256u;
        for (uint32_t base = blockIdx.x * 256u + wave * 64u; base < nmov; base += stride) {
            const uint32_t i = base + lane;
            const bool valid = i < nmov;
            const uint32_t key = valid ? mv.key[i] : 0u;
            const uint32_t slot = valid ? mv.slot[i] : 0u;
            const uint64_t heads = run_heads(key, valid);
            uint32_t hv = SLOSH_NONE;
            if (valid && ((heads >> lane) & 1ull)) hv = hash_find(keys, cap, key);
            const int head = run_head_of(heads);
            const uint32_t hid = __shfl(hv, head);
            const bool ok = valid && hid != SLOSH_NONE;
            if (valid && hid == SLOSH_NONE) atomicOr(&counters[CTR_ERRORS], ERR_LOOKUP);
            uint32_t r0 = 0u;
            if (ok && ((heads >> lane) & 1ull)) r0 = atomicAdd(&mv.mcount[hid], run_length(heads));
            const uint32_t r = __shfl(r0, head) + (lane - (uint32_t)head);
            const bool over = ok && r >= MOVER_CAP;
            if (ok && !over) mv.mlist[(size_t)hid * MOVER_CAP + r] = slot;
            const uint64_t om = __ballot(over);
            if (om) {
                const int first = __ffsll((unsigned long long)om) - 1;
                uint32_t ob = 0u;
                if ((int)lane == first) ob = atomicAdd(&counters[CTR_OVERFLOW], (uint32_t)__popcll(om));
                ob = __shfl(ob, first);
                if (over) {
                    const uint32_t q = ob + (uint32_t)__popcll(om & lanemask_lt());
                    mv.ovf_slot[q] = slot;
                    mv.ovf_hid[q] = hid;
                }
            }
        }
        return;
    }
    if (blockIdx.x < mv_wgs + cmp_wgs) {
        const uint32_t w = blockIdx.x - mv_wgs;
        constexpr uint32_t SLOTS = 256u * ITEMS;
        for (uint32_t base = w * SLOTS; base < cap; base += cmp_wgs * SLOTS) {
            uint32_t occ[ITEMS];
            uint32_t mine = 0;
#pragma unroll
            for (uint32_t i = 0; i < ITEMS; i++) {
                const uint32_t s = base + tid * ITEMS + i;
                occ[i] = s < cap && keys[s] != SLOSH_NONE ? 1u : 0u;
                mine += occ[i];
            }
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
            for (uint32_t q = 0; q < 4; q++) {
                before += q < wave ? wsum[q] : 0u;
                total += wsum[q];
            }
            if (tid == 0) wg_base = total ? atomicAdd(&counters[CTR_NUM_BLOCKS], total) : 0u;
            __syncthreads();
            uint32_t o = wg_base + before + x - mine;
#pragma unroll
            for (uint32_t i = 0; i < ITEMS; i++)
                if (occ[i]) blist[o++] = base + tid * ITEMS + i;
            __syncthreads();  // wsum / wg_base are reused by the next run
        }
        return;
    }
    const uint32_t nbr_wgs = gridDim.x - mv_wgs - cmp_wgs;
    for (uint32_t s = (blockIdx.x - mv_wgs - cmp_wgs) * 16u + (tid >> 4); s < cap; s += nbr_wgs * 16u) {
        const uint32_t key = keys[s];
        if (key != SLOSH_NONE) block_neighbour(s, key, tid & 15u, nbr, keys, prev_keys, cap, batched, mv);
    }
}

// update_block_particle_count (sort.slang:208-246) of the full sort: block
// totals and per-(block, cell) ranks.
__global__ __launch_bounds__(256) void k_count(const float4 *__restrict__ pos, uint32_t n, float h, uint32_t batched,
                                               const uint32_t *__restrict__ keys, uint32_t cap,
                                               uint32_t *__restrict__ cell_cnt, uint32_t *__restrict__ pslot,
                                               uint32_t *__restrict__ prank, uint32_t *__restrict__ counters,
                                               uint32_t *__restrict__ blk_cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
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
    const uint64_t bad = __ballot(!in_range);
    if (bad && lane == (uint32_t)(__ffsll((unsigned long long)bad) - 1)) atomicOr(&counters[CTR_ERRORS], ERR_KEY_RANGE);
    // Hash lookup by the head of each run of equal block keys, all runs at once.
    const uint64_t bheads = run_heads(key, valid);
    uint32_t hv = SLOSH_NONE;
    if (valid && ((bheads >> lane) & 1ull)) hv = hash_find(keys, cap, key);
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

// Per block (one wavefront): exclusive scan of the 64 cell counts (offsets
// relative to the block's first slot). Waves [w0, w0 + nw) of the grid take
// list entries w0, w0 + nw, ...
__device__ inline void cell_starts(const uint32_t *__restrict__ blist, const uint32_t *__restrict__ cell_cnt,
                                   uint32_t *__restrict__ cell_start, uint32_t B, uint32_t w, uint32_t nw) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t k = w; k < B; k += nw) {
        const uint32_t hid = blist[k];
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

// Exclusive scan of v[base, base + 4096) by one workgroup of 1024 threads (4
// consecutive entries per thread), written through `out` with `carry` added;
// returns the tile's total. Entry e is v[e] (list == null) or the block total
// blk_cnt[list[e]] written to out[list[e]]. Every thread must call it.
__device__ inline uint32_t scan_tile(const uint32_t *__restrict__ list, const uint32_t *__restrict__ v,
                                     uint32_t *__restrict__ out, uint32_t base, uint32_t B, uint32_t carry,
                                     uint32_t *wsum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t idx = base + tid * 4u;
    uint32_t at[4], val[4];
#pragma unroll
    for (int j = 0; j < 4; j++) at[j] = idx + j < B ? (list ? list[idx + j] : idx + j) : 0u;
#pragma unroll
    for (int j = 0; j < 4; j++) val[j] = idx + j < B ? v[at[j]] : 0u;
    const uint32_t t = val[0] + val[1] + val[2] + val[3];
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
    uint32_t e = carry + (wave > 0 ? wsum[wave - 1] : 0u) + x - t;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (idx + j < B) out[at[j]] = e;
        e += val[j];
    }
    const uint32_t total = wsum[15];
    __syncthreads();
    return total;
}

// Exclusive scan of the block totals in block-list order, in one workgroup
// (used while the capacity is at most a few tiles). Workgroups 1.. compute the
// per-block cell offsets meanwhile (independent work in the same launch).
__global__ __launch_bounds__(1024) void k_block_scan(const uint32_t *__restrict__ blist,
                                                     const uint32_t *__restrict__ blk_cnt,
                                                     uint32_t *__restrict__ blk_start,
                                                     uint32_t *__restrict__ counters,
                                                     const uint32_t *__restrict__ cell_cnt,
                                                     uint32_t *__restrict__ cell_start) {
    __shared__ uint32_t wsum[16];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    if (blockIdx.x > 0) {
        cell_starts(blist, cell_cnt, cell_start, B, (blockIdx.x - 1) * 16u + (threadIdx.x >> 6),
                    (gridDim.x - 1) * 16u);
        return;
    }
    uint32_t carry = 0;
    for (uint32_t base = 0; base < B; base += 4096u) carry += scan_tile(blist, blk_cnt, blk_start, base, B, carry, wsum);
    if (threadIdx.x == 0) counters[CTR_SORTED_ALLOC] = carry;
}

// Three-phase scan for large block counts: tiles of 4096 list entries scanned
// in parallel, then the tile totals, then the tile offsets added.
__global__ __launch_bounds__(1024) void k_block_scan_tiles(const uint32_t *__restrict__ blist,
                                                           const uint32_t *__restrict__ blk_cnt,
                                                           uint32_t *__restrict__ blk_start,
                                                           uint32_t *__restrict__ tile_sums,
                                                           const uint32_t *__restrict__ counters,
                                                           uint32_t scan_wgs, const uint32_t *__restrict__ cell_cnt,
                                                           uint32_t *__restrict__ cell_start) {
    __shared__ uint32_t wsum[16];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    if (blockIdx.x >= scan_wgs) {  // the remaining workgroups compute the cell offsets
        cell_starts(blist, cell_cnt, cell_start, B, (blockIdx.x - scan_wgs) * 16u + (threadIdx.x >> 6),
                    (gridDim.x - scan_wgs) * 16u);
        return;
    }
    for (uint32_t t = blockIdx.x; t * 4096u < B; t += scan_wgs) {
        const uint32_t total = scan_tile(blist, blk_cnt, blk_start, t * 4096u, B, 0u, wsum);
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
        carry += scan_tile(nullptr, tile_sums, tile_sums, base, ntiles, carry, wsum);
    if (threadIdx.x == 0) counters[CTR_SORTED_ALLOC] = carry;
}

__global__ __launch_bounds__(256) void k_block_scan_add(const uint32_t *__restrict__ blist,
                                                        uint32_t *__restrict__ blk_start,
                                                        const uint32_t *__restrict__ tile_sums,
                                                        const uint32_t *__restrict__ counters) {
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    for (uint32_t k = 4096u + blockIdx.x * blockDim.x + threadIdx.x; k < B; k += gridDim.x * blockDim.x)
        blk_start[blist[k]] += tile_sums[k >> 12];
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

static inline uint32_t div_up(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }
static inline uint32_t min_u(uint32_t a, uint32_t b) { return a < b ? a : b; }

// `fast`: the previous substep's fused G2P kernel filled this table, wrote every
// block's stayers first and listed the movers; one launch then sorts. Otherwise
// the full sort. `ev` (optional) brackets the whole sequence.
hipError_t launch_grid_build(const GridBuffers &g, const float4 *pos, uint32_t n, float h, uint32_t batched,
                             uint32_t cap, bool fast, const PassEvents *ev, hipStream_t stream) {
    const uint32_t slot_grid = min_u(div_up(cap, 256), 2048);
    const uint32_t part_grid = div_up(n, 256);
    MoverLists mv;
    mv.key = g.mv_key;
    mv.slot = g.mv_slot;
    mv.mcount = g.mcount;
    mv.mlist = g.mlist;
    mv.ovf_slot = g.ovf_slot;
    mv.ovf_hid = g.ovf_hid;
    mv.prev_start = fast ? g.prev_start : nullptr;
    mv.prev_cnt = fast ? g.prev_cnt : nullptr;
    mv.prev_stay = fast ? g.prev_stay : nullptr;
    mv.bin = fast ? g.bin : nullptr;
    const bool wide = cap > 65536u;
    const uint32_t cmp_wgs = wide ? min_u(div_up(cap, 2048), 2048) : min_u(div_up(cap, 256), 2048);
    const uint32_t nbr_wgs = min_u(div_up(cap, 16), 4096);
    auto sort_kernel = wide ? k_sort_fast<8> : k_sort_fast<1>;
    if (fast) {
        // Grid-stride over the movers (their count is on the device).
        const uint32_t mv_wgs = n > 0 ? min_u(div_up(n, 4096), 1024) : 0u;
        launch_kernel(ev, sort_kernel, dim3(mv_wgs + cmp_wgs + nbr_wgs), dim3(256), 0, stream,
                      (const uint32_t *)g.keys, (const uint32_t *)g.prev_keys, cap, batched, g.counters, g.blist,
                      g.nbr, mv_wgs, cmp_wgs, mv);
        return hipGetLastError();
    }
    PassEvents first, last;
    if (ev) {
        first.start = ev->start;
        last.stop = ev->stop;
    }
    launch_kernel(ev ? &first : nullptr, k_reset_table, dim3(slot_grid), dim3(256), 0, stream, g.keys, cap,
                  g.counters);
    if (n > 0) {
        BlockInit init;
        init.mcount = g.mcount;
        init.blk_cnt = g.blk_cnt;
        init.cell_cnt = g.cell_cnt;
        launch_kernel(nullptr, k_touch_primary, dim3(part_grid), dim3(256), 0, stream, pos, n, h, batched, g.keys,
                      cap, g.counters, init);
    }
    const PassEvents *sort_ev = n > 0 ? nullptr : (ev ? &last : nullptr);
    launch_kernel(sort_ev, sort_kernel, dim3(cmp_wgs + nbr_wgs), dim3(256), 0, stream, (const uint32_t *)g.keys,
                  (const uint32_t *)g.prev_keys, cap, batched, g.counters, g.blist, g.nbr, 0u, cmp_wgs, mv);
    if (n == 0) return hipGetLastError();
    launch_kernel(nullptr, k_count, dim3(part_grid), dim3(256), 0, stream, pos, n, h, batched,
                  (const uint32_t *)g.keys, cap, g.cell_cnt, g.pslot, g.prank, g.counters, g.blk_cnt);
    const uint32_t cell_wgs = min_u(div_up(cap, 16), 1024);  // 16 blocks (wavefronts) per workgroup
    if (cap <= 16384u) {
        launch_kernel(nullptr, k_block_scan, dim3(1 + cell_wgs), dim3(1024), 0, stream, (const uint32_t *)g.blist,
                      (const uint32_t *)g.blk_cnt, g.blk_start, g.counters, (const uint32_t *)g.cell_cnt,
                      g.cell_start);
    } else {
        const uint32_t scan_wgs = min_u(div_up(cap, 4096), 1024);
        launch_kernel(nullptr, k_block_scan_tiles, dim3(scan_wgs + cell_wgs), dim3(1024), 0, stream,
                      (const uint32_t *)g.blist, (const uint32_t *)g.blk_cnt, g.blk_start, g.tile_sums,
                      (const uint32_t *)g.counters, scan_wgs, (const uint32_t *)g.cell_cnt, g.cell_start);
        launch_kernel(nullptr, k_block_scan_top, dim3(1), dim3(1024), 0, stream, g.tile_sums, g.counters);
        launch_kernel(nullptr, k_block_scan_add, dim3(min_u(div_up(cap, 256), 4096)), dim3(256), 0, stream,
                      (const uint32_t *)g.blist, g.blk_start, (const uint32_t *)g.tile_sums,
                      (const uint32_t *)g.counters);
    }
    launch_kernel(ev ? &last : nullptr, k_scatter, dim3(part_grid), dim3(256), 0, stream, (const uint32_t *)g.pslot,
                  (const uint32_t *)g.prank, (const uint32_t *)g.cell_start, (const uint32_t *)g.blk_start, g.perm, n);
    return hipGetLastError();
}
