// Device-side particle storage maintenance between substeps: removal of whole
// scenes or of an id range with order-preserving compaction, and the gather of
// one scene's particles in creation order. The reference does these on the host
// by re-uploading tensors (GpuParticles::shift_remove / append,
// src/solver/particle.rs:505-572); here the particles never leave HBM.
//
//   k_mark_keep        keep flag per storage slot and per original id
//   k_scan_tile / k_scan_tiles / k_scan_add
//                      three-phase exclusive scan of u32 flags (2048 per tile)
//   k_compact_slots    copies kept slots to the other particle buffer at their
//                      scanned position and renumbers the original ids
//   k_compact_static   compacts the by-id material records
//   k_scene_gather     writes one scene's particles, in id order, to a staging
//                      buffer set
#include "../../include/slosh_hip.h"
#include "launch.h"

namespace {

constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ inline uint32_t slot_scene(float4 p) { return __float_as_uint(p.w) >> 1; }

// A particle is removed if its scene is flagged in `dead` (mode 0) or its id lies
// in [lo, hi) (mode 1).
__device__ inline bool removed(uint32_t mode, const uint8_t *__restrict__ dead, uint32_t scene, uint32_t id,
                               uint32_t lo, uint32_t hi) {
    if (mode == 0) return scene < SLOSH_MAX_SCENES && dead[scene] != 0;
    return id >= lo && id < hi;
}

__global__ __launch_bounds__(256) void k_mark_keep(const float4 *__restrict__ pos, const uint32_t *__restrict__ ids,
                                                   uint32_t n, uint32_t mode, const uint8_t *__restrict__ dead,
                                                   uint32_t lo, uint32_t hi, uint32_t *__restrict__ keep_slot,
                                                   uint32_t *__restrict__ keep_id, uint32_t *__restrict__ keep_id2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    const uint32_t k = removed(mode, dead, slot_scene(pos[i]), id, lo, hi) ? 0u : 1u;
    keep_slot[i] = k;
    if (id < n) {
        keep_id[id] = k;
        keep_id2[id] = k;
    }
}

// Exclusive wave scan of one value per lane; returns the prefix, *total = sum.
__device__ inline uint32_t wave_exclusive(uint32_t v, uint32_t *total) {
    const uint32_t lane = lane_id();
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// Phase 1: each workgroup scans its tile in place and writes the tile total.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile(uint32_t *__restrict__ data, uint32_t n,
                                                           uint32_t *__restrict__ tile_sums) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const uint32_t base = blockIdx.x * SCAN_TILE + tid * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? data[base + k] : 0u;
        s += v[k];
    }
    uint32_t wt;
    const uint32_t pre = wave_exclusive(s, &wt);
    if (lane_id() == 63) wsum[wave] = wt;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < SCAN_THREADS / 64; w++) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    uint32_t run = off + pre;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
    if (tid == 0) tile_sums[blockIdx.x] = tot;
}

// Phase 2: one workgroup scans the tile totals; total_out receives the sum.
__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t *__restrict__ tile_sums, uint32_t ntiles,
                                                    uint32_t *__restrict__ total_out) {
    __shared__ uint32_t wsum[16];
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 1024u) {
        const uint32_t i = base + tid;
        const uint32_t v = i < ntiles ? tile_sums[i] : 0u;
        uint32_t wt;
        const uint32_t pre = wave_exclusive(v, &wt);
        if (lane_id() == 63) wsum[wave] = wt;
        __syncthreads();
        uint32_t off = 0, tot = 0;
#pragma unroll
        for (uint32_t w = 0; w < 16; w++) {
            if (w < wave) off += wsum[w];
            tot += wsum[w];
        }
        if (i < ntiles) tile_sums[i] = carry + off + pre;
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) *total_out = carry;
}

// Phase 3: adds each tile's offset.
__global__ __launch_bounds__(256) void k_scan_add(uint32_t *__restrict__ data, uint32_t n,
                                                 const uint32_t *__restrict__ tile_sums) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] += tile_sums[i / SCAN_TILE];
}

struct PSet {
    float4 *pos, *vm;
    float *C, *F, *vgd, *dps, *fdt;
    uint32_t *ids;
};

__global__ __launch_bounds__(256) void k_compact_slots(PSet in, PSet out, uint32_t n, uint32_t pcap, uint32_t mode,
                                                       const uint8_t *__restrict__ dead, uint32_t lo, uint32_t hi,
                                                       const uint32_t *__restrict__ new_slot,
                                                       const uint32_t *__restrict__ new_id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = in.pos[i];
    const uint32_t id = in.ids[i];
    if (removed(mode, dead, slot_scene(p), id, lo, hi)) return;
    const uint32_t j = new_slot[i];
    out.pos[j] = p;
    out.vm[j] = in.vm[i];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        out.C[(size_t)k * pcap + j] = in.C[(size_t)k * pcap + i];
        out.F[(size_t)k * pcap + j] = in.F[(size_t)k * pcap + i];
    }
    out.vgd[j] = in.vgd[i];
    if (in.dps)
#pragma unroll
        for (int k = 0; k < 3; k++) out.dps[(size_t)k * pcap + j] = in.dps[(size_t)k * pcap + i];
#pragma unroll
    for (int k = 0; k < 3; k++) out.fdt[(size_t)k * pcap + j] = in.fdt[(size_t)k * pcap + i];
    out.ids[j] = new_id[id];
}

struct SSet {
    float4 *prop, *model, *dp0, *dp1;
};

__global__ __launch_bounds__(256) void k_compact_static(SSet in, SSet out, uint32_t n,
                                                        const uint32_t *__restrict__ keep_id,
                                                        const uint32_t *__restrict__ new_id) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n || !keep_id[id]) return;
    const uint32_t j = new_id[id];
    out.prop[j] = in.prop[id];
    out.model[j] = in.model[id];
    if (in.dp0) {
        out.dp0[j] = in.dp0[id];
        out.dp1[j] = in.dp1[id];
    }
}

__global__ __launch_bounds__(256) void k_mark_scene(const float4 *__restrict__ pos, const uint32_t *__restrict__ ids,
                                                    uint32_t n, uint32_t scene, uint32_t *__restrict__ flag_id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    if (id < n) flag_id[id] = slot_scene(pos[i]) == scene ? 1u : 0u;
}

// Staging layout: pos/vm as float4 [count], C/F as [9][count], vgd/fdt/dps as
// [count] / [3][count].
__global__ __launch_bounds__(256) void k_scene_gather(PSet in, PSet out, uint32_t n, uint32_t pcap, uint32_t scene,
                                                      uint32_t count, const uint32_t *__restrict__ rank_id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = in.pos[i];
    if (slot_scene(p) != scene) return;
    const uint32_t j = rank_id[in.ids[i]];
    if (j >= count) return;
    out.pos[j] = p;
    out.vm[j] = in.vm[i];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        out.C[(size_t)k * count + j] = in.C[(size_t)k * pcap + i];
        out.F[(size_t)k * count + j] = in.F[(size_t)k * pcap + i];
    }
    out.vgd[j] = in.vgd[i];
#pragma unroll
    for (int k = 0; k < 3; k++) out.fdt[(size_t)k * count + j] = in.fdt[(size_t)k * pcap + i];
    if (in.dps)
#pragma unroll
        for (int k = 0; k < 3; k++) out.dps[(size_t)k * count + j] = in.dps[(size_t)k * pcap + i];
}

inline uint32_t div_up_u(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

// In-place exclusive scan of data[0, n); *total (device) receives the sum.
// tile_sums holds at least scan_tiles(n) words.
hipError_t scan_u32(uint32_t *data, uint32_t n, uint32_t *tile_sums, uint32_t *total, hipStream_t st) {
    const uint32_t tiles = div_up_u(n > 0 ? n : 1u, SCAN_TILE);
    hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(SCAN_THREADS), 0, st, data, n, tile_sums);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tile_sums, tiles, total);
    if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3(div_up_u(n, 256)), dim3(256), 0, st, data, n, (const uint32_t *)tile_sums);
    return hipGetLastError();
}

}  // namespace

uint32_t scan_tiles(uint32_t n) { return div_up_u(n > 0 ? n : 1u, SCAN_TILE); }

hipError_t launch_compact(const CompactArgs &a, uint32_t n, uint32_t pcap, uint32_t mode, const uint8_t *dead,
                          uint32_t lo, uint32_t hi, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t g = div_up_u(n, 256);
    hipLaunchKernelGGL(k_mark_keep, dim3(g), dim3(256), 0, st, (const float4 *)a.pos[0], (const uint32_t *)a.ids[0],
                       n, mode, dead, lo, hi, a.keep_slot, a.keep_id, a.keep_id2);
    hipError_t e = scan_u32(a.keep_slot, n, a.tile_sums, a.total, st);
    if (e != hipSuccess) return e;
    e = scan_u32(a.keep_id, n, a.tile_sums, a.total, st);
    if (e != hipSuccess) return e;
    PSet in{a.pos[0], a.vm[0], a.C[0], a.F[0], a.vgd[0], a.dps[0], a.fdt[0], a.ids[0]};
    PSet out{a.pos[1], a.vm[1], a.C[1], a.F[1], a.vgd[1], a.dps[1], a.fdt[1], a.ids[1]};
    hipLaunchKernelGGL(k_compact_slots, dim3(g), dim3(256), 0, st, in, out, n, pcap, mode, dead, lo, hi,
                       (const uint32_t *)a.keep_slot, (const uint32_t *)a.keep_id);
    SSet sin{a.prop[0], a.model[0], a.dp0[0], a.dp1[0]};
    SSet sout{a.prop[1], a.model[1], a.dp0[1], a.dp1[1]};
    hipLaunchKernelGGL(k_compact_static, dim3(g), dim3(256), 0, st, sin, sout, n, (const uint32_t *)a.keep_id2,
                       (const uint32_t *)a.keep_id);
    return hipGetLastError();
}

hipError_t launch_scene_gather(const GatherArgs &a, uint32_t n, uint32_t pcap, uint32_t scene, uint32_t count,
                               hipStream_t st) {
    if (n == 0 || count == 0) return hipSuccess;
    const uint32_t g = div_up_u(n, 256);
    hipLaunchKernelGGL(k_mark_scene, dim3(g), dim3(256), 0, st, (const float4 *)a.pos, (const uint32_t *)a.ids, n,
                       scene, a.rank_id);
    hipError_t e = scan_u32(a.rank_id, n, a.tile_sums, a.total, st);
    if (e != hipSuccess) return e;
    PSet in{a.pos, a.vm, a.C, a.F, a.vgd, a.dps, a.fdt, a.ids};
    PSet out{a.s_pos, a.s_vm, a.s_C, a.s_F, a.s_vgd, a.s_dps, a.s_fdt, nullptr};
    hipLaunchKernelGGL(k_scene_gather, dim3(g), dim3(256), 0, st, in, out, n, pcap, scene, count,
                       (const uint32_t *)a.rank_id);
    return hipGetLastError();
}
