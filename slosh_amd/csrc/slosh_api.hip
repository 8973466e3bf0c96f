// Host implementation of the C ABI declared in include/slosh_hip.h.
//
// Device state layout (one slosh_data):
//   particles, two buffers (current / next), indexed by storage slot:
//     pos   float4  xyz, w = (scene << 1 | enabled) as u32 bits
//     vm    float4  velocity xyz, mass
//     C     float[9][pcap]  APIC affine matrix (velocity gradient between G2P
//                           and the particle update), shader matrix row-major
//     F     float[9][pcap]  deformation gradient
//     vgd   float[pcap]     vel_grad_det
//     ids   u32[pcap]       original particle index of the slot
//     dps   float[3][pcap]  Drucker-Prager state (allocated iff a sand particle exists)
//   force_dt float[3][pcap] by current slot (zero after every particle update)
//   static records by original index: prop (V0, damping, fixed, radius),
//     model (tag, lambda, mu, cfl of the elastic part), dp0 (h0..h3), dp1 (lambda, mu)
//   grid: two hash tables (ping-pong = previous substep; a block's header id is
//     the slot of its key), per block of each table its storage range, stayer
//     count and mover count, and per block the block list entry, neighbour
//     table, 64 cell counts/starts, MOVER_CAP mover slots, a 6x6x6 float4 P2G
//     tile, 64 float4 nodes, two collision caches (moving colliders).
// The fused G2P/particle-update kernel writes the next buffer block by block
// (each block's stayers first, its leavers last, listed for the next sort), so
// storage follows the block order of the previous substep; readback restores
// the original order through `ids`.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/slosh_hip.h"
#include "launch.h"

// ---- error plumbing ---------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return fail(SLOSH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(SLOSH_E_NO_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    return SLOSH_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

#define TRY(expr)                    \
    do {                             \
        int r_ = (expr);             \
        if (r_ != SLOSH_OK) return r_; \
    } while (0)

// ---- objects ----------------------------------------------------------------
struct slosh_pipeline {
    slosh_kernels kernels;
};

struct PBufs {
    float4 *pos = nullptr, *vm = nullptr;
    float *C = nullptr, *F = nullptr, *vgd = nullptr, *dps = nullptr;
    uint32_t *ids = nullptr;
};

// Creation-time records of consecutive particles of one scene. A run whose
// particles all share one model and one property record (a scene of one
// material) keeps that record once; `expand` turns it into one record per
// particle before per-particle edits.
struct HostRun {
    uint32_t scene = 0;
    size_t count = 0;
    bool uniform = false;
    std::vector<slosh_particle_model> recs_m;
    std::vector<slosh_particle_properties> recs_p;
    const slosh_particle_model &model(size_t i) const { return uniform ? recs_m[0] : recs_m[i]; }
    const slosh_particle_properties &props(size_t i) const { return uniform ? recs_p[0] : recs_p[i]; }
    void expand() {
        if (!uniform) return;
        const slosh_particle_model m0 = recs_m[0];  // assign() must not take a reference into the vector
        const slosh_particle_properties p0 = recs_p[0];
        recs_m.assign(count, m0);
        recs_p.assign(count, p0);
        uniform = false;
    }
};

// A user-captured substep sequence. It is registered with its data: destroying
// or mutating the data retires the executable, so a launch can never replay
// kernels over freed or reallocated buffers.
struct slosh_graph {
    slosh_data *data = nullptr;  // null once the data was destroyed
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;   // recorded after each launch
    uint64_t generation = 0;
    int cur_p = 0, cur_g = 0;
    StepParams params;           // step parameters the substeps were captured with
};

struct slosh_data {
    slosh_pipeline *pipeline = nullptr;
    int device = -1;  // the device current at creation, which holds every buffer
    size_t n = 0, pcap = 0;
    uint32_t cap = 0;
    float h = 0.0f;
    slosh_sim_params params{};
    uint32_t batched = 0;
    uint32_t has_sand = 0;
    uint32_t has_force = 0;
    uint32_t coll_cache = 0;
    uint32_t coll_prev_valid = 0;
    uint32_t svd_f64 = 1;
    uint32_t tag_mask = 0;  // material tags present (bit t = tag t, bit 4 = other)
    uint32_t uniform = 0;   // all particles share one static material record
    float4 u_prop{}, u_model{}, u_dp0{}, u_dp1{};
    int cur_p = 0;  // current particle buffer
    int cur_g = 0;  // hash table / collision cache written by the last substep
    // The table of the next substep was reset and filled with every particle's
    // primary block by the last substep's fused kernels, which also wrote every
    // block's stayers first and listed the movers: the next substep sorts in one
    // launch.
    bool next_primed = false;
    PBufs pb[2];
    float *fdt = nullptr;
    float4 *s_prop = nullptr, *s_model = nullptr, *s_dp0 = nullptr, *s_dp1 = nullptr;
    // destinations of the device-side compaction (swapped with the above)
    float *fdt2 = nullptr;
    float4 *s2_prop = nullptr, *s2_model = nullptr, *s2_dp0 = nullptr, *s2_dp1 = nullptr;
    uint32_t *scan_tmp = nullptr;  // scan tile sums + total
    uint8_t *dead = nullptr;       // per-scene removal flags [SLOSH_MAX_SCENES]
    std::vector<uint8_t> dead_host;  // source of the last flag upload (outlives the async copy)
    slosh_tuning tuning{512u, 256u, 8192u};
    std::vector<slosh_boundary_condition> mats_host;
    // Creation-time records in original-id order, as runs of consecutive ids
    // with one scene id (an appended scene is one run): removing a scene drops
    // runs instead of rewriting per-particle host arrays.
    std::vector<HostRun> runs;
    uint64_t tag_count[5] = {0, 0, 0, 0, 0};  // particles per material tag (4: other)
    int uni_state = 0;                        // 0: no particles, 1: uniform (u_* valid), 2: mixed
    // Batched data: per scene id, 0 no particle given, 1 one static record (its
    // first, in scene_first), 2 several; scenes_mixed counts the 2s. While it is
    // zero the kernels read each scene's record from scene_rec (filled on the
    // device from the uploaded by-id records) instead of gathering by id.
    std::vector<uint8_t> scene_state;
    std::vector<float4> scene_first;
    uint32_t scenes_mixed = 0;
    // Per scene id, as mass_state per scene: 0 no particle given, 1 every particle
    // given with mass = enabled ? scene_mass_bits : +0 in one upload, 2 otherwise;
    // scene_masses_mixed counts the 2s. scene_masses_off once masses were
    // rewritten by the caller.
    std::vector<uint8_t> scene_mass_state;
    std::vector<uint32_t> scene_mass_bits;
    uint32_t scene_masses_mixed = 0;
    bool scene_masses_off = false;
    float4 *scene_rec = nullptr;  // [4 * SLOSH_MAX_SCENES]: prop, model, dp0, dp1
    // Particle masses: 0 no particles, 1 every particle created enabled with mass
    // u_mass (bitwise) and not exposed to hooks or device views since, 2 otherwise.
    int mass_state = 0;
    float u_mass = 0.0f;
    bool masses_exposed = false;  // a hook ran or a device view was handed out
    // grid: two hash tables (a block's header id is its key's slot) and, per
    // table, each block's storage range, stayer count and received movers
    uint32_t *keys[2] = {nullptr, nullptr};
    uint32_t *blk_start[2] = {nullptr, nullptr}, *blk_cnt[2] = {nullptr, nullptr};
    uint32_t *stay[2] = {nullptr, nullptr}, *mcount[2] = {nullptr, nullptr};
    uint32_t *blist = nullptr, *nbr = nullptr, *cell_cnt = nullptr, *cell_start = nullptr;
    uint32_t *mlist = nullptr;  // MOVER_CAP movers per block
    uint4 *bin = nullptr;       // per block: its previous-table block's range and stayers
    uint32_t *pslot = nullptr, *prank = nullptr, *perm = nullptr;
    // per particle: movers listed for the next substep, and movers beyond a
    // block's list slots
    uint32_t *mv_key = nullptr, *mv_slot = nullptr, *ovf_slot = nullptr, *ovf_hid = nullptr;
    uint32_t *blk_tile_sums = nullptr;
    uint32_t *dt_bound = nullptr;  // GpuTimestepBounds word
    uint32_t *counters = nullptr;
    float4 *tiles = nullptr, *nodes = nullptr;
    float4 *coll[2] = {nullptr, nullptr};
    // bodies
    uint32_t nbodies = 0;
    std::vector<slosh_body_desc> bodies_host;
    float4 *shapes = nullptr, *poses = nullptr, *vels = nullptr, *mprops = nullptr, *local_com = nullptr;
    // Trimesh collision buffers (slosh_data_set_collision_mesh); has_trimesh: some
    // body is a trimesh, so the buffers are required.
    float *mesh_vtx = nullptr;
    uint32_t *mesh_idx = nullptr;
    size_t mesh_nv = 0, mesh_ni = 0;
    bool has_trimesh = false;
    // Pinned host arenas that appends stage particles in, each reused once the
    // uploads queued from it (marked by `done`) have completed. Up to 32, so that
    // a burst of admissions while substeps run does not wait for them: the uploads
    // of an arena in use are queued behind those substeps.
    struct StageArena {
        uint8_t *ptr = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
    };
    static constexpr int MAX_STAGE_ARENAS = 32;
    StageArena stage[MAX_STAGE_ARENAS];
    int num_stage = 0;
    // A pinned ring reserved ahead of time (slosh_reserve_staging): an append that
    // fits takes the region after the last one, so admissions pin nothing. Regions
    // are freed in order, once the uploads queued from them (marked by `done`) have
    // completed.
    struct RingRegion {
        size_t off, len;
        hipEvent_t done;
    };
    uint8_t *ring = nullptr;
    size_t ring_bytes = 0, ring_head = 0;
    std::deque<RingRegion> ring_used;
    std::vector<hipEvent_t> ring_events;  // idle events for new regions
    uint2 *mats = nullptr;
    int *impulses = nullptr;
    // Captured substep pairs, one per starting parity, each with the step
    // parameters it was captured with (a replay under other parameters is
    // recaptured) and an event recorded after its last launch.
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    StepParams graph_params[4];
    hipEvent_t graph_done[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<slosh_graph *> user_graphs;  // live slosh_graph objects of this data
    uint64_t generation = 0;  // bumped by every mutation; user graphs check it
    hipStream_t capture_stream = nullptr;
};

struct slosh_timestamps {
    struct Rec {
        int pass;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double acc[SLOSH_NUM_PASSES] = {0};
    uint32_t count[SLOSH_NUM_PASSES] = {0};
};

enum { PASS_GRID_SORT = 0, PASS_P2G, PASS_GRID_UPDATE, PASS_G2P, PASS_PARTICLES_UPDATE, PASS_INTEGRATE_BODIES };

// ---- helpers ----------------------------------------------------------------
static uint32_t next_pow2(uint32_t v) {
    uint32_t c = 1;
    while (c < v) c <<= 1;
    return c;
}

static float u2f(uint32_t x) {
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}
static uint32_t f2u(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    return x;
}

// Two alternating sets of per-substep counters, one per hash table.
static uint32_t *ctr_set(slosh_data *d, int g) { return d->counters + 8 * g; }

// Releases a graph executable once its last launch has completed: a replay may
// still be running on the GPU, and the runtime frees the executable's kernel
// arguments with it.
static void retire_exec(hipGraphExec_t &exec, hipEvent_t done) {
    if (!exec) return;
    if (done) (void)hipEventSynchronize(done);
    (void)hipGraphExecDestroy(exec);
    exec = nullptr;
}

// Records `*ev` (created on first use) on `st` after a graph launch, unless `st`
// is being captured (then the launch is a node of the caller's graph).
static void mark_launch(hipEvent_t *ev, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    if (!*ev && hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) {
        *ev = nullptr;
        (void)hipStreamSynchronize(st);  // no event: make the launch complete instead
        return;
    }
    (void)hipEventRecord(*ev, st);
}

// Every change to the state a captured substep reads from its kernel arguments
// (buffers, counts, step parameters, uniform records, mass tracking, collision
// cache) goes through here: cached pairs and user graphs are retired.
static void invalidate_graphs(slosh_data *d) {
    d->next_primed = false;
    d->generation++;
    for (int i = 0; i < 4; i++) retire_exec(d->graph[i], d->graph_done[i]);
    for (slosh_graph *g : d->user_graphs) retire_exec(g->exec, g->done);
}

static int alloc_particle_buffers(slosh_data *d, size_t pcap, bool sand) {
    for (int b = 0; b < 2; b++) {
        PBufs &p = d->pb[b];
        TRY(dalloc(&p.pos, pcap));
        TRY(dalloc(&p.vm, pcap));
        TRY(dalloc(&p.C, 9 * pcap));
        TRY(dalloc(&p.F, 9 * pcap));
        TRY(dalloc(&p.vgd, pcap));
        TRY(dalloc(&p.ids, pcap));
        if (sand) TRY(dalloc(&p.dps, 3 * pcap));
    }
    TRY(dalloc(&d->fdt, 3 * pcap));
    TRY(dalloc(&d->fdt2, 3 * pcap));
    TRY(dalloc(&d->s_prop, pcap));
    TRY(dalloc(&d->s_model, pcap));
    TRY(dalloc(&d->s2_prop, pcap));
    TRY(dalloc(&d->s2_model, pcap));
    if (sand) {
        TRY(dalloc(&d->s_dp0, pcap));
        TRY(dalloc(&d->s_dp1, pcap));
        TRY(dalloc(&d->s2_dp0, pcap));
        TRY(dalloc(&d->s2_dp1, pcap));
    }
    TRY(dalloc(&d->scan_tmp, scan_tiles((uint32_t)pcap) + 4));
    TRY(dalloc(&d->pslot, pcap));
    TRY(dalloc(&d->prank, pcap));
    TRY(dalloc(&d->perm, pcap));
    TRY(dalloc(&d->mv_key, pcap));
    TRY(dalloc(&d->mv_slot, pcap));
    TRY(dalloc(&d->ovf_slot, pcap));
    TRY(dalloc(&d->ovf_hid, pcap));
    // Every entry a kernel can read is a particle index even before the first sort.
    HIP_TRY(hipMemset(d->perm, 0, sizeof(uint32_t) * (pcap ? pcap : 1)));
    HIP_TRY(hipMemset(d->fdt, 0, sizeof(float) * 3 * (pcap ? pcap : 1)));
    return SLOSH_OK;
}

static void free_particle_buffers(slosh_data *d) {
    for (int b = 0; b < 2; b++) {
        PBufs &p = d->pb[b];
        dfree(p.pos);
        dfree(p.vm);
        dfree(p.C);
        dfree(p.F);
        dfree(p.vgd);
        dfree(p.ids);
        dfree(p.dps);
    }
    dfree(d->fdt);
    dfree(d->s_prop);
    dfree(d->s_model);
    dfree(d->s_dp0);
    dfree(d->s_dp1);
    dfree(d->fdt2);
    dfree(d->s2_prop);
    dfree(d->s2_model);
    dfree(d->s2_dp0);
    dfree(d->s2_dp1);
    dfree(d->scan_tmp);
    dfree(d->pslot);
    dfree(d->prank);
    dfree(d->perm);
    dfree(d->mv_key);
    dfree(d->mv_slot);
    dfree(d->ovf_slot);
    dfree(d->ovf_hid);
}

static bool model_is_sand(const slosh_particle_model &m) { return m.words[0] == 2u || m.words[0] == 3u; }

// Static per-particle records as the kernels read them: prop (V0, damping, fixed,
// radius), model (tag, elastic lambda, mu, cfl), dp0 (h0..h3), dp1 (plastic lambda, mu).
static void static_records(const slosh_particle_properties &pr, const slosh_particle_model &m, float4 *prop,
                           float4 *model, float4 *dp0, float4 *dp1);

static int tag_slot(const slosh_particle_model &m) { return m.words[0] < 4u ? (int)m.words[0] : 4; }

// d->uniform is set while every particle has the same static material record
// (the kernels then take it from their arguments instead of gathering it);
// tracked incrementally as particles are added, reset when the data empties.
static void note_records(slosh_data *d, const slosh_particle_model *m, const slosh_particle_properties *pr, size_t n,
                         int sign) {
    for (size_t i = 0; i < n; i++) d->tag_count[tag_slot(m[i])] += (uint64_t)(int64_t)sign;
    uint32_t mask = 0;
    for (int t = 0; t < 5; t++)
        if (d->tag_count[t]) mask |= 1u << t;
    d->tag_mask = mask;
    if (sign > 0) {
        for (size_t i = 0; i < n && d->uni_state != 2; i++) {
            float4 p, mm, a, b;
            static_records(pr[i], m[i], &p, &mm, &a, &b);
            if (d->uni_state == 0) {
                d->u_prop = p;
                d->u_model = mm;
                d->u_dp0 = a;
                d->u_dp1 = b;
                d->uni_state = 1;
            } else if (std::memcmp(&p, &d->u_prop, 12) != 0 || std::memcmp(&mm, &d->u_model, 16) != 0 ||
                       std::memcmp(&a, &d->u_dp0, 16) != 0 || std::memcmp(&b, &d->u_dp1, 8) != 0) {
                d->uni_state = 2;
            }
        }
    }
    if (d->n == 0 && sign < 0) {
        d->uni_state = 0;
        if (!d->masses_exposed) d->mass_state = 0;
    }
    d->uniform = d->uni_state == 1 ? 1u : 0u;
}

// note_records(-1) for particles [a, b) of run r.
static void unnote_run(slosh_data *d, const HostRun &r, size_t a, size_t b) {
    if (r.uniform) {
        d->tag_count[tag_slot(r.recs_m[0])] -= (uint64_t)(b - a);
        note_records(d, nullptr, nullptr, 0, -1);  // refreshes the tag mask
    } else {
        note_records(d, r.recs_m.data() + a, r.recs_p.data() + a, b - a, -1);
    }
}

// Tracks whether every particle given so far has mass = enabled ? u_mass : +0
// (bitwise), the form the fused update keeps (it zeroes the mass it disables).
static void note_masses(slosh_data *d, const slosh_kinematics *k, size_t n) {
    for (size_t i = 0; i < n && d->mass_state != 2; i++) {
        if (!k[i].enabled) {
            if (f2u(k[i].mass) != 0u) d->mass_state = 2;
        } else if (d->mass_state == 0) {
            d->u_mass = k[i].mass;
            d->mass_state = 1;
        } else if (f2u(k[i].mass) != f2u(d->u_mass)) {
            d->mass_state = 2;
        }
    }
}

// Per-scene record tracking (batched data): the static records of one particle
// of scene `sc`, compared as note_records compares them.
static void note_scene(slosh_data *d, uint32_t sc, const slosh_particle_model &m, const slosh_particle_properties &pr) {
    if (d->scene_state.empty()) {
        d->scene_state.assign(SLOSH_MAX_SCENES, 0);
        d->scene_first.assign(4 * (size_t)SLOSH_MAX_SCENES, float4{});
    }
    if (d->scene_state[sc] == 2) return;
    float4 r[4];
    static_records(pr, m, &r[0], &r[1], &r[2], &r[3]);
    float4 *f = &d->scene_first[4 * (size_t)sc];
    if (d->scene_state[sc] == 0) {
        for (int k = 0; k < 4; k++) f[k] = r[k];
        d->scene_state[sc] = 1;
    } else if (std::memcmp(&r[0], &f[0], 12) != 0 || std::memcmp(&r[1], &f[1], 16) != 0 ||
               std::memcmp(&r[2], &f[2], 16) != 0 || std::memcmp(&r[3], &f[3], 8) != 0) {
        d->scene_state[sc] = 2;
        d->scenes_mixed++;
    }
}

// Per-scene mass tracking of one upload (batched data). States: 0 no particle
// given, 1 only disabled particles with mass +0, 3 also enabled ones, all with
// mass scene_mass_bits, 2 anything else. A scene given particles by a second
// upload counts as 2: the device takes a scene's mass from its latest upload.
static void scene_mass_mixed(slosh_data *d, uint32_t sc) {
    if (d->scene_mass_state[sc] != 2) {
        d->scene_mass_state[sc] = 2;
        d->scene_masses_mixed++;
    }
}

static void note_scene_masses(slosh_data *d, const slosh_particles *p) {
    if (d->scene_mass_state.empty()) {
        d->scene_mass_state.assign(SLOSH_MAX_SCENES, 0);
        d->scene_mass_bits.assign(SLOSH_MAX_SCENES, 0u);
    }
    std::vector<uint32_t> mine;  // the scenes this upload gives particles to
    uint32_t last = SLOSH_NONE;
    for (size_t i = 0; i < p->len; i++) {
        const uint32_t sc = p->scene_ids[i];
        if (sc != last) {
            last = sc;
            if (std::find(mine.begin(), mine.end(), sc) == mine.end()) {
                mine.push_back(sc);
                if (d->scene_mass_state[sc] != 0) scene_mass_mixed(d, sc);
            }
        }
        uint8_t &st = d->scene_mass_state[sc];
        if (st == 2) continue;
        const slosh_kinematics &k = p->kinematics[i];
        const uint32_t bits = f2u(k.mass);
        if (!k.enabled) {
            if (bits != 0u) scene_mass_mixed(d, sc);
            else if (st == 0) st = 1;
        } else if (st == 3) {
            if (bits != d->scene_mass_bits[sc]) scene_mass_mixed(d, sc);
        } else {
            d->scene_mass_bits[sc] = bits;
            st = 3;
        }
    }
}

// A removed scene's id may be reused with other records.
static void forget_scene(slosh_data *d, uint32_t sc) {
    if (!d->scene_state.empty()) {
        if (d->scene_state[sc] == 2) d->scenes_mixed--;
        d->scene_state[sc] = 0;
    }
    if (!d->scene_mass_state.empty()) {
        if (d->scene_mass_state[sc] == 2) d->scene_masses_mixed--;
        d->scene_mass_state[sc] = 0;
        d->scene_mass_bits[sc] = 0u;
    }
}

static bool same_records(const slosh_particle_model &a, const slosh_particle_properties &pa,
                         const slosh_particle_model &b, const slosh_particle_properties &pb) {
    return std::memcmp(&a, &b, sizeof a) == 0 && std::memcmp(&pa, &pb, sizeof pa) == 0;
}

// Appends records to the host runs (split where the scene id changes). A range
// whose records are all equal joins or starts a uniform run without copying them.
static void host_append(slosh_data *d, const slosh_particles *p) {
    size_t i = 0;
    while (i < p->len) {
        const uint32_t sc = p->scene_ids ? p->scene_ids[i] : 0u;
        size_t j = i + 1;
        while (j < p->len && (p->scene_ids ? p->scene_ids[j] : 0u) == sc) j++;
        if (d->runs.empty() || d->runs.back().scene != sc) {
            d->runs.emplace_back();
            d->runs.back().scene = sc;
        }
        HostRun &r = d->runs.back();
        bool uni = true;
        for (size_t k = i + 1; k < j && uni; k++)
            uni = same_records(p->models[k], p->properties[k], p->models[i], p->properties[i]);
        if (d->batched) {
            if (uni) note_scene(d, sc, p->models[i], p->properties[i]);
            else
                for (size_t k = i; k < j; k++) note_scene(d, sc, p->models[k], p->properties[k]);
        }
        if (uni && r.count == 0) {
            r.recs_m.assign(1, p->models[i]);
            r.recs_p.assign(1, p->properties[i]);
            r.uniform = true;
        } else if (!(uni && r.uniform && same_records(r.recs_m[0], r.recs_p[0], p->models[i], p->properties[i]))) {
            r.expand();
            r.recs_m.insert(r.recs_m.end(), p->models + i, p->models + j);
            r.recs_p.insert(r.recs_p.end(), p->properties + i, p->properties + j);
        }
        r.count += j - i;
        i = j;
    }
}

static std::vector<slosh_particle_model> flat_models(const slosh_data *d) {
    std::vector<slosh_particle_model> v;
    v.reserve(d->n);
    for (const auto &r : d->runs) {
        if (r.uniform) v.insert(v.end(), r.count, r.recs_m[0]);
        else v.insert(v.end(), r.recs_m.begin(), r.recs_m.end());
    }
    return v;
}
static std::vector<slosh_particle_properties> flat_props(const slosh_data *d) {
    std::vector<slosh_particle_properties> v;
    v.reserve(d->n);
    for (const auto &r : d->runs) {
        if (r.uniform) v.insert(v.end(), r.count, r.recs_p[0]);
        else v.insert(v.end(), r.recs_p.begin(), r.recs_p.end());
    }
    return v;
}
static std::vector<uint32_t> flat_scenes(const slosh_data *d) {
    std::vector<uint32_t> v;
    v.reserve(d->n);
    for (const auto &r : d->runs) v.insert(v.end(), r.count, r.scene);
    return v;
}
// Writes models (original-id order) back into the runs.
static void set_models(slosh_data *d, const std::vector<slosh_particle_model> &m) {
    size_t o = 0;
    for (auto &r : d->runs) {
        r.expand();
        std::copy(m.begin() + o, m.begin() + o + r.count, r.recs_m.begin());
        o += r.count;
    }
}

// Host-side staging of particles in the internal layout (SoA, slot order).
struct Staging {
    size_t n = 0;
    std::vector<float> pos, vm, C, F, vgd, dps, fdt;
    std::vector<uint32_t> ids;
    std::vector<float> prop, model, dp0, dp1;  // 4 floats per particle (by original index)
    void resize(size_t count) {
        n = count;
        pos.assign(4 * count, 0.f);
        vm.assign(4 * count, 0.f);
        C.assign(9 * count, 0.f);
        F.assign(9 * count, 0.f);
        vgd.assign(count, 0.f);
        dps.assign(3 * count, 0.f);
        fdt.assign(3 * count, 0.f);
        ids.assign(count, 0u);
        prop.assign(4 * count, 0.f);
        model.assign(4 * count, 0.f);
        dp0.assign(4 * count, 0.f);
        dp1.assign(4 * count, 0.f);
    }
};

// The same arrays carved from a pinned host arena (50 words per particle): host
// to device copies from it run at the DMA rate and need not be waited for. The
// arrays written as float4 come first (16 n bytes each), so each starts 16-byte
// aligned for any count; the scalar planes follow.
struct PinnedStaging {
    static constexpr size_t WORDS = 50;
    size_t n = 0;
    float *pos, *vm, *C, *F, *vgd, *dps, *fdt, *prop, *model, *dp0, *dp1;
    uint32_t *ids;
    PinnedStaging(uint8_t *arena, size_t count) : n(count) {
        float *w = reinterpret_cast<float *>(arena);
        pos = w; w += 4 * count;
        vm = w; w += 4 * count;
        prop = w; w += 4 * count;
        model = w; w += 4 * count;
        dp0 = w; w += 4 * count;
        dp1 = w; w += 4 * count;
        C = w; w += 9 * count;
        F = w; w += 9 * count;
        vgd = w; w += count;
        dps = w; w += 3 * count;
        fdt = w; w += 3 * count;
        ids = reinterpret_cast<uint32_t *>(w);
    }
    void resize(size_t) {}
};

// Stages particles [lo, hi) of `p`; flags[0..2] = any force, any sand, bad scene id.
template <typename S>
static void stage_range(const slosh_particles *p, uint32_t id_base, S &s, size_t lo, size_t hi, uint8_t *flags) {
    const size_t n = p->len;
    for (size_t i = lo; i < hi; i++) {
        const slosh_kinematics &k = p->kinematics[i];
        const slosh_particle_model &m = p->models[i];
        const slosh_particle_properties &pr = p->properties[i];
        const uint32_t scene = p->scene_ids ? p->scene_ids[i] : 0u;
        if (scene >= SLOSH_MAX_SCENES) flags[2] = 1;
        s.pos[4 * i + 0] = p->positions[4 * i + 0];
        s.pos[4 * i + 1] = p->positions[4 * i + 1];
        s.pos[4 * i + 2] = p->positions[4 * i + 2];
        s.pos[4 * i + 3] = u2f((scene << 1) | (k.enabled != 0u ? 1u : 0u));
        for (int c = 0; c < 3; c++) s.vm[4 * i + c] = k.velocity[c];
        s.vm[4 * i + 3] = k.mass;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                s.C[(size_t)(r * 3 + c) * n + i] = k.affine.m[r][c];
                s.F[(size_t)(r * 3 + c) * n + i] = p->def_grad[i].m[r][c];
            }
        s.vgd[i] = k.vel_grad_det;
        s.ids[i] = id_base + (uint32_t)i;
        for (int c = 0; c < 3; c++) {
            s.fdt[(size_t)c * n + i] = k.force_dt[c];
            if (k.force_dt[c] != 0.0f) flags[0] = 1;
        }
        for (int c = 0; c < 4; c++) s.dp0[4 * i + c] = s.dp1[4 * i + c] = 0.0f;
        static_records(pr, m, reinterpret_cast<float4 *>(&s.prop[4 * i]), reinterpret_cast<float4 *>(&s.model[4 * i]),
                       reinterpret_cast<float4 *>(&s.dp0[4 * i]), reinterpret_cast<float4 *>(&s.dp1[4 * i]));
        const bool sand = model_is_sand(m);
        if (sand) flags[1] = 1;
        for (int c = 0; c < 3; c++) s.dps[(size_t)c * n + i] = sand ? u2f(m.words[1 + c]) : 0.0f;
    }
}

// AoS reference records -> the internal SoA staging. The loop scatters each
// particle into ~20 arrays (~40 ns per particle on one core), so large inputs
// are split over up to 8 host threads; a range whose thread cannot be started
// runs on the caller's.
// Host threads for a pass over n particles (one below 2^17 particles).
static size_t host_threads(size_t n) {
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency() / 2);
    return n < ((size_t)1 << 17) ? 1 : std::min<size_t>(std::min<size_t>(hw, 8), n >> 16);
}

// One pass over an append's particles (in parallel for large ones): whether they
// all belong to the first one's scene and share its records, and how their
// masses relate (enabled ones' bits all equal, disabled ones' all +0).
struct AppendSummary {
    bool one_scene = true, one_record = true;
    bool has_enabled = false, enabled_equal = true, disabled_nonzero = false;
    uint32_t mass_bits = 0u;
};

static AppendSummary summarize_range(const slosh_particles *p, size_t lo, size_t hi) {
    AppendSummary a;
    const uint32_t sc0 = p->scene_ids ? p->scene_ids[0] : 0u;
    for (size_t i = lo; i < hi; i++) {
        if (p->scene_ids && p->scene_ids[i] != sc0) a.one_scene = false;
        if (a.one_record && !same_records(p->models[i], p->properties[i], p->models[0], p->properties[0]))
            a.one_record = false;
        const slosh_kinematics &k = p->kinematics[i];
        const uint32_t bits = f2u(k.mass);
        if (!k.enabled) {
            if (bits != 0u) a.disabled_nonzero = true;
        } else if (!a.has_enabled) {
            a.has_enabled = true;
            a.mass_bits = bits;
        } else if (bits != a.mass_bits) {
            a.enabled_equal = false;
        }
    }
    return a;
}

static AppendSummary summarize(const slosh_particles *p) {
    const size_t n = p->len, T = host_threads(n);
    std::vector<AppendSummary> part(T);
    std::vector<std::thread> workers;
    for (size_t t = 1; t < T; t++) {
        try {
            workers.emplace_back([&, t]() { part[t] = summarize_range(p, n * t / T, n * (t + 1) / T); });
        } catch (...) {
            part[t] = summarize_range(p, n * t / T, n * (t + 1) / T);
        }
    }
    part[0] = summarize_range(p, 0, n / T);
    for (auto &w : workers) w.join();
    AppendSummary a;
    for (const AppendSummary &b : part) {
        a.one_scene = a.one_scene && b.one_scene;
        a.one_record = a.one_record && b.one_record;
        a.disabled_nonzero = a.disabled_nonzero || b.disabled_nonzero;
        a.enabled_equal = a.enabled_equal && b.enabled_equal;
        if (b.has_enabled) {
            if (!a.has_enabled) {
                a.has_enabled = true;
                a.mass_bits = b.mass_bits;
            } else if (b.mass_bits != a.mass_bits) {
                a.enabled_equal = false;
            }
        }
    }
    return a;
}

// host_append + note_records(+1) + note_masses (+ note_scene_masses) for one
// append: for the common case (one scene, one record) from the summary, with
// the same results as the per-particle passes, which run otherwise.
static void note_append(slosh_data *d, const slosh_particles *p) {
    if (p->len == 0) return;
    const AppendSummary a = summarize(p);
    const uint32_t sc = p->scene_ids ? p->scene_ids[0] : 0u;
    HostRun *last = d->runs.empty() ? nullptr : &d->runs.back();
    const bool joins = last && last->scene == sc && last->count > 0;
    const bool fast = a.one_scene && a.one_record &&
                      (!joins || (last->uniform && same_records(last->recs_m[0], last->recs_p[0], p->models[0],
                                                                p->properties[0])));
    if (!fast) {
        host_append(d, p);
        note_records(d, p->models, p->properties, p->len, 1);
        note_masses(d, p->kinematics, p->len);
        if (d->batched) note_scene_masses(d, p);
        return;
    }
    // host_append of one uniform range
    if (!last || last->scene != sc) {
        d->runs.emplace_back();
        d->runs.back().scene = sc;
    }
    HostRun &r = d->runs.back();
    if (r.count == 0) {
        r.recs_m.assign(1, p->models[0]);
        r.recs_p.assign(1, p->properties[0]);
        r.uniform = true;
    }
    if (d->batched) note_scene(d, sc, p->models[0], p->properties[0]);
    r.count += p->len;
    // note_records: every particle adds to its tag's count; the records are the first one's
    d->tag_count[tag_slot(p->models[0])] += (uint64_t)p->len - 1u;
    note_records(d, p->models, p->properties, 1, 1);
    // note_masses
    if (d->mass_state != 2) {
        if (a.disabled_nonzero || !a.enabled_equal) {
            d->mass_state = 2;
        } else if (a.has_enabled) {
            if (d->mass_state == 0) {
                d->u_mass = u2f(a.mass_bits);
                d->mass_state = 1;
            } else if (a.mass_bits != f2u(d->u_mass)) {
                d->mass_state = 2;
            }
        }
    }
    // note_scene_masses
    if (d->batched) {
        if (d->scene_mass_state.empty()) {
            d->scene_mass_state.assign(SLOSH_MAX_SCENES, 0);
            d->scene_mass_bits.assign(SLOSH_MAX_SCENES, 0u);
        }
        if (d->scene_mass_state[sc] != 0 || a.disabled_nonzero || !a.enabled_equal) {
            scene_mass_mixed(d, sc);
        } else if (a.has_enabled) {
            d->scene_mass_bits[sc] = a.mass_bits;
            d->scene_mass_state[sc] = 3;
        } else {
            d->scene_mass_state[sc] = 1;
        }
    }
}

template <typename S>
static int stage_particles(const slosh_particles *p, uint32_t id_base, S &s, bool *any_force, bool *any_sand) {
    const size_t n = p->len;
    s.resize(n);
    const size_t T = host_threads(n);
    std::vector<uint8_t> flags(3 * T, 0);
    std::vector<std::thread> workers;
    std::vector<size_t> inline_ranges;
    for (size_t t = 1; t < T; t++) {
        try {
            workers.emplace_back(stage_range<S>, p, id_base, std::ref(s), n * t / T, n * (t + 1) / T, &flags[3 * t]);
        } catch (...) {
            inline_ranges.push_back(t);
        }
    }
    stage_range<S>(p, id_base, s, 0, n / T, &flags[0]);
    for (size_t t : inline_ranges) stage_range<S>(p, id_base, s, n * t / T, n * (t + 1) / T, &flags[3 * t]);
    for (auto &w : workers) w.join();
    for (size_t t = 0; t < T; t++) {
        if (flags[3 * t + 2]) return fail(SLOSH_E_INVALID, "scene id out of range");
        if (flags[3 * t]) *any_force = true;
        if (flags[3 * t + 1]) *any_sand = true;
    }
    return SLOSH_OK;
}

static void static_records(const slosh_particle_properties &pr, const slosh_particle_model &m, float4 *prop,
                           float4 *model, float4 *dp0, float4 *dp1) {
    *prop = make_float4(pr.init_volume, pr.damping, u2f(pr.fixed), pr.init_radius);
    const uint32_t tag = m.words[0];
    *model = make_float4(u2f(tag), 0.0f, 0.0f, 0.0f);
    *dp0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    *dp1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tag == 0u || tag == 1u) {
        *model = make_float4(u2f(tag), u2f(m.words[1]), u2f(m.words[2]), u2f(m.words[3]));
    } else if (tag == 2u || tag == 3u) {
        *model = make_float4(u2f(tag), u2f(m.words[10]), u2f(m.words[11]), u2f(m.words[12]));
        *dp0 = make_float4(u2f(m.words[4]), u2f(m.words[5]), u2f(m.words[6]), u2f(m.words[7]));
        *dp1 = make_float4(u2f(m.words[8]), u2f(m.words[9]), 0.0f, 0.0f);
    }
}

// Uploads staged particles into slots [slot0, slot0+n) of buffer `b` (ids starting at id0).
// With `wait`, returns once the copies are done (the staging may then be freed);
// otherwise the copies stay queued on `st` ahead of whatever follows.
template <typename S>
static int upload_staging(slosh_data *d, const S &s, int b, size_t slot0, size_t id0, hipStream_t st,
                          bool wait = true) {
    const size_t n = s.n;
    if (n == 0) return SLOSH_OK;
    PBufs &p = d->pb[b];
    const size_t pc = d->pcap;
    HIP_TRY(hipMemcpyAsync(p.pos + slot0, &s.pos[0], 16 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p.vm + slot0, &s.vm[0], 16 * n, hipMemcpyHostToDevice, st));
    for (int k = 0; k < 9; k++) {
        HIP_TRY(hipMemcpyAsync(p.C + k * pc + slot0, &s.C[k * n], 4 * n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(p.F + k * pc + slot0, &s.F[k * n], 4 * n, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(p.vgd + slot0, &s.vgd[0], 4 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p.ids + slot0, &s.ids[0], 4 * n, hipMemcpyHostToDevice, st));
    if (p.dps)
        for (int k = 0; k < 3; k++)
            HIP_TRY(hipMemcpyAsync(p.dps + k * pc + slot0, &s.dps[k * n], 4 * n, hipMemcpyHostToDevice, st));
    for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(d->fdt + k * pc + slot0, &s.fdt[k * n], 4 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d->s_prop + id0, &s.prop[0], 16 * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d->s_model + id0, &s.model[0], 16 * n, hipMemcpyHostToDevice, st));
    if (d->s_dp0) {
        HIP_TRY(hipMemcpyAsync(d->s_dp0 + id0, &s.dp0[0], 16 * n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d->s_dp1 + id0, &s.dp1[0], 16 * n, hipMemcpyHostToDevice, st));
    }
    if (d->scene_rec) {
        StaticBuffers sb{d->s_prop, d->s_model, d->s_dp0, d->s_dp1, nullptr};
        HIP_TRY(launch_scene_records(p.pos, p.vm, p.ids, (uint32_t)slot0, (uint32_t)n, sb, d->scene_rec, st));
    }
    if (wait) HIP_TRY(hipStreamSynchronize(st));
    return SLOSH_OK;
}

// Downloads the current particle buffer (slot order).
struct Snapshot {
    std::vector<float> pos, vm, C, F, vgd, dps, fdt;
    std::vector<uint32_t> ids;
};

static int download_current(slosh_data *d, hipStream_t st, Snapshot &s) {
    const size_t n = d->n, pc = d->pcap;
    PBufs &p = d->pb[d->cur_p];
    s.pos.resize(4 * n);
    s.vm.resize(4 * n);
    s.C.resize(9 * n);
    s.F.resize(9 * n);
    s.vgd.resize(n);
    s.ids.resize(n);
    s.fdt.resize(3 * n);
    s.dps.assign(3 * n, 0.f);
    if (n == 0) return SLOSH_OK;
    HIP_TRY(hipMemcpyAsync(s.pos.data(), p.pos, 16 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s.vm.data(), p.vm, 16 * n, hipMemcpyDeviceToHost, st));
    for (int k = 0; k < 9; k++) {
        HIP_TRY(hipMemcpyAsync(s.C.data() + k * n, p.C + k * pc, 4 * n, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s.F.data() + k * n, p.F + k * pc, 4 * n, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpyAsync(s.vgd.data(), p.vgd, 4 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s.ids.data(), p.ids, 4 * n, hipMemcpyDeviceToHost, st));
    if (p.dps)
        for (int k = 0; k < 3; k++)
            HIP_TRY(hipMemcpyAsync(s.dps.data() + k * n, p.dps + k * pc, 4 * n, hipMemcpyDeviceToHost, st));
    for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(s.fdt.data() + k * n, d->fdt + k * pc, 4 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; i++)
        if (s.ids[i] >= n) return fail(SLOSH_E_GRID_OVERFLOW, "particle storage inconsistent (grid overflow?)");
    return SLOSH_OK;
}

static StepParams make_params(const slosh_data *d) {
    StepParams P;
    std::memset(&P, 0, sizeof P);  // padding included: captured parameter sets compare bytewise
    P.gravity[0] = d->params.gravity[0];
    P.gravity[1] = d->params.gravity[1];
    P.gravity[2] = d->params.gravity[2];
    P.dt = d->params.dt;
    P.h = d->h;
    P.inv_h = 1.0f / d->h;
    P.h_dt = d->h / d->params.dt;
    P.inv_d = 4.0f / (d->h * d->h);
    P.cap = d->cap;
    P.n = (uint32_t)d->n;
    P.batched = d->batched;
    P.num_bodies = d->nbodies;
    P.has_force = d->has_force;
    P.has_sand = d->has_sand;
    P.coll_cache = d->coll_cache;
    P.cur = (uint32_t)d->cur_g;
    P.svd_f64 = d->svd_f64;
    P.tag_mask = d->tag_mask;
    P.uniform = d->uniform;
    P.scene_uniform = (d->batched && d->scene_rec && d->scenes_mixed == 0 && !d->uniform) ? 1u : 0u;
    P.scene_mass = (d->batched && d->scene_rec && !d->masses_exposed && !d->scene_masses_off &&
                    d->scene_masses_mixed == 0 && d->mass_state != 1)
                       ? 1u
                       : 0u;
    P.p2g_chunk = d->tuning.p2g_chunk;
    P.g2p_threads = d->tuning.g2p_threads;
    P.block_wgs = d->tuning.block_wgs;
    P.u_prop = d->u_prop;
    P.u_model = d->u_model;
    P.u_dp0 = d->u_dp0;
    P.u_dp1 = d->u_dp1;
    P.mass_uniform = d->mass_state == 1 ? 1u : 0u;
    P.u_mass = d->u_mass;
    return P;
}

static ParticleBuffers pbufs(slosh_data *d, int b) {
    ParticleBuffers r;
    r.pos = d->pb[b].pos;
    r.vm = d->pb[b].vm;
    r.C = d->pb[b].C;
    r.F = d->pb[b].F;
    r.vgd = d->pb[b].vgd;
    r.dps = d->pb[b].dps;
    r.ids = d->pb[b].ids;
    return r;
}

static StaticBuffers sbufs(slosh_data *d) {
    StaticBuffers s;
    s.prop = d->s_prop;
    s.model = d->s_model;
    s.dp0 = d->s_dp0;
    s.dp1 = d->s_dp1;
    s.scene_rec = d->scene_rec;
    return s;
}

// Events bracketing one pass (written by the dispatch packets), or nullptr.
static const PassEvents *pass_events(slosh_timestamps *ts, int pass, PassEvents *storage) {
    if (!ts) return nullptr;
    slosh_timestamps::Rec r;
    r.pass = pass;
    if (ts->pool.size() >= 2) {
        r.a = ts->pool.back();
        ts->pool.pop_back();
        r.b = ts->pool.back();
        ts->pool.pop_back();
    } else {
        (void)hipEventCreateWithFlags(&r.a, hipEventDisableSystemFence);
        (void)hipEventCreateWithFlags(&r.b, hipEventDisableSystemFence);
    }
    ts->recs.push_back(r);
    storage->start = r.a;
    storage->stop = r.b;
    return storage;
}

// ---- one substep --------------------------------------------------------------
static int run_step(slosh_pipeline *pl, slosh_data *d, hipStream_t st, const slosh_hooks *hooks, void *hs,
                    slosh_timestamps *ts) {
    if (d->has_trimesh && !d->mesh_vtx)
        return fail(SLOSH_E_INVALID, "trimesh colliders need slosh_data_set_collision_mesh before stepping");
    const slosh_kernels &K = pl->kernels;
    const bool any_hook = hooks && (hooks->after_particle_sort || hooks->run_p2g || hooks->run_g2p || hooks->after_p2g ||
                                    hooks->after_grid_update || hooks->after_g2p || hooks->after_particles_update);
    // Hooks may write particle masses through their device views. Captured
    // substeps carry the mass tracking and the collision-cache switch in their
    // arguments, so changing either retires them.
    if (any_hook && (!d->masses_exposed || d->mass_state != 2)) {
        d->masses_exposed = true;
        d->mass_state = 2;
        invalidate_graphs(d);
    }
    if (any_hook && d->nbodies > 0 && !d->coll_cache) {
        // Hooks may move bodies: from now on keep the reference's per-node cache.
        for (int i = 0; i < 2; i++)
            if (!d->coll[i]) TRY(dalloc(&d->coll[i], 2 * (size_t)d->cap * NODES_PER_BLOCK));
        d->coll_cache = 1;
        d->coll_prev_valid = 0;
        invalidate_graphs(d);
    }
    // GpuGrid::swap_buffers (grid.rs:513-517): the table written last substep becomes "previous".
    const int prev_g = d->cur_g;
    const int cur_g = prev_g ^ 1;
    d->cur_g = cur_g;
    StepParams P = make_params(d);
    const uint32_t n = (uint32_t)d->n;

    // Steady state: the previous substep's fused kernels prepared this table and
    // the mover list, so the sort is one launch. Hooks observe or replace passes
    // with the full sort's result (the sorted list), so they get the full sort.
    const bool fast = d->next_primed && !any_hook;
    GridBuffers g;
    g.keys = d->keys[cur_g];
    g.prev_keys = d->keys[prev_g];
    g.blist = d->blist;
    g.nbr = d->nbr;
    g.cell_cnt = d->cell_cnt;
    g.cell_start = d->cell_start;
    g.blk_start = d->blk_start[cur_g];
    g.blk_cnt = d->blk_cnt[cur_g];
    g.pslot = d->pslot;
    g.prank = d->prank;
    g.perm = d->perm;
    g.counters = ctr_set(d, cur_g);
    g.tile_sums = d->blk_tile_sums;
    g.mv_key = d->mv_key;
    g.mv_slot = d->mv_slot;
    g.mcount = d->mcount[cur_g];
    g.mlist = d->mlist;
    g.ovf_slot = d->ovf_slot;
    g.ovf_hid = d->ovf_hid;
    g.prev_start = d->blk_start[prev_g];
    g.prev_cnt = d->blk_cnt[prev_g];
    g.prev_stay = d->stay[prev_g];
    g.bin = d->bin;

    // The fused path (no hooks, built-in P2G and particle update) prepares the
    // next substep's table and sort inside this substep's kernels.
    const bool prime = !any_hook && K.builtin_transfers && K.particles_update;
    NextTableBuffers next;
    next.keys = d->keys[prev_g];
    next.counters = ctr_set(d, prev_g);
    next.mcount = d->mcount[prev_g];
    next.mv_key = d->mv_key;
    next.mv_slot = d->mv_slot;
    next.prime = prime ? 1u : 0u;

    PassEvents pe;
    HIP_TRY(launch_grid_build(g, d->pb[d->cur_p].pos, n, d->h, d->batched, d->cap, fast,
                              pass_events(ts, PASS_GRID_SORT, &pe), st));
    d->next_primed = false;
    if (hooks && hooks->after_particle_sort && hooks->after_particle_sort(d, st, hs) != 0)
        return fail(SLOSH_E_HOOK, "after_particle_sort hook failed");

    TransferBuffers t;
    t.fast = fast ? 1u : 0u;
    t.keys = d->keys[cur_g];
    t.blist = d->blist;
    t.nbr = d->nbr;
    t.counters = ctr_set(d, cur_g);
    t.errors = ctr_set(d, cur_g) + CTR_ERRORS;
    t.perm = d->perm;
    t.cell_cnt = d->cell_cnt;
    t.blk_start = d->blk_start[cur_g];
    t.blk_cnt = d->blk_cnt[cur_g];
    t.stay = d->stay[cur_g];
    t.bin = d->bin;
    t.mcount = d->mcount[cur_g];
    t.mlist = d->mlist;
    t.ovf_slot = d->ovf_slot;
    t.ovf_hid = d->ovf_hid;
    t.tiles = d->tiles;
    t.nodes = d->nodes;
    t.coll_prev = (d->coll_cache && d->coll_prev_valid) ? (void *)d->coll[prev_g] : nullptr;
    t.coll_cur = d->coll_cache ? (void *)d->coll[cur_g] : nullptr;
    t.fdt = d->fdt;

    int handled = 0;
    if (hooks && hooks->run_p2g && hooks->run_p2g(d, st, hs, &handled) != 0)
        return fail(SLOSH_E_HOOK, "run_p2g hook failed");
    const bool p2g_by_hook = handled != 0;
    if (!p2g_by_hook) {
        if (!K.builtin_transfers)
            return fail(SLOSH_E_NO_KERNELS, "pipeline built without builtin transfer kernels; a hook must handle run_p2g");
        HIP_TRY(launch_p2g(P, pbufs(d, d->cur_p), (uint32_t)d->pcap, t, next, pass_events(ts, PASS_P2G, &pe), st));
    }
    if (hooks && hooks->after_p2g && hooks->after_p2g(d, st, hs) != 0) return fail(SLOSH_E_HOOK, "after_p2g hook failed");

    ColliderBuffers cb;
    cb.shapes = d->shapes;
    cb.poses = d->poses;
    cb.mats = d->mats;
    cb.count = d->nbodies;
    cb.mesh_vtx = d->has_trimesh ? (const float4 *)d->mesh_vtx : nullptr;
    cb.mesh_idx = d->has_trimesh ? d->mesh_idx : nullptr;
    if (K.grid_update || !p2g_by_hook) {
        const uint32_t mode = p2g_by_hook ? 1u : 0u;
        HIP_TRY(launch_grid_update(P, t, cb, mode, K.grid_update ? 1u : 0u, pass_events(ts, PASS_GRID_UPDATE, &pe),
                                   st));
    }
    if (d->coll_cache) d->coll_prev_valid = 1;
    if (hooks && hooks->after_grid_update && hooks->after_grid_update(d, st, hs) != 0)
        return fail(SLOSH_E_HOOK, "after_grid_update hook failed");

    if (p2g_by_hook) next.prime = 0u;  // the table of the next substep was not reset
    BodyBuffers bb;
    bb.poses = d->poses;
    bb.vels = d->vels;
    bb.mprops = d->mprops;
    bb.local_com = d->local_com;
    bb.impulses = d->impulses;
    bb.count = d->nbodies;
    handled = 0;
    if (hooks && hooks->run_g2p && hooks->run_g2p(d, st, hs, &handled) != 0)
        return fail(SLOSH_E_HOOK, "run_g2p hook failed");
    const bool g2p_by_hook = handled != 0;
    const bool split = g2p_by_hook || (hooks && hooks->after_g2p) || !K.particles_update;
    if (!g2p_by_hook && !K.builtin_transfers)
        return fail(SLOSH_E_NO_KERNELS, "pipeline built without builtin transfer kernels; a hook must handle run_g2p");
    if (!split) {
        // Fused G2P + particle update; writes the next particle buffer in sorted order.
        // Fused G2P + particle update; writes the next particle buffer in sorted
        // order and (fused) integrates the rigid bodies.
        HIP_TRY(launch_g2p_update(P, pbufs(d, d->cur_p), pbufs(d, d->cur_p ^ 1), sbufs(d), (uint32_t)d->pcap, t, next,
                                  bb, pass_events(ts, PASS_G2P, &pe), st));
        d->cur_p ^= 1;
        d->next_primed = next.prime != 0u;
    } else {
        if (!g2p_by_hook) {
            HIP_TRY(launch_g2p(P, pbufs(d, d->cur_p), (uint32_t)d->pcap, t, pass_events(ts, PASS_G2P, &pe), st));
        }
        if (hooks && hooks->after_g2p && hooks->after_g2p(d, st, hs) != 0)
            return fail(SLOSH_E_HOOK, "after_g2p hook failed");
        if (K.particles_update) {
            HIP_TRY(launch_particle_update(P, pbufs(d, d->cur_p), sbufs(d), (uint32_t)d->pcap, d->fdt,
                                           pass_events(ts, PASS_PARTICLES_UPDATE, &pe), st));
        }
    }
    if (K.particles_update) d->has_force = 0;  // force_dt is zeroed by every particle update
    if (hooks && hooks->after_particles_update && hooks->after_particles_update(d, st, hs) != 0)
        return fail(SLOSH_E_HOOK, "after_particles_update hook failed");

    if (split && d->nbodies > 0) {
        HIP_TRY(launch_rigid_update(P, bb, pass_events(ts, PASS_INTEGRATE_BODIES, &pe), st));
    }
    return SLOSH_OK;
}

// ---- C ABI --------------------------------------------------------------------
extern "C" {

const char *slosh_last_error(void) { return g_last_error.c_str(); }
const char *slosh_version(void) { return "slosh_hip 0.1 (gfx950)"; }

void slosh_kernels_default(slosh_kernels *out) {
    if (!out) return;
    out->builtin_transfers = 1;
    out->cdf = 1;
    out->rigid_particles = 1;
    out->node_reset = 1;
    out->grid_update = 1;
    out->particles_update = 1;
}

int slosh_pipeline_create(const slosh_kernels *kernels, slosh_pipeline **out) {
    if (!out) return fail(SLOSH_E_INVALID, "null output pointer");
    slosh_pipeline *p = new slosh_pipeline();
    if (kernels) p->kernels = *kernels;
    else slosh_kernels_default(&p->kernels);
    *out = p;
    return SLOSH_OK;
}

void slosh_pipeline_destroy(slosh_pipeline *p) { delete p; }

// Waits for every region of the staging ring and frees it.
static void release_ring(slosh_data *d) {
    for (auto &r : d->ring_used) {
        (void)hipEventSynchronize(r.done);
        d->ring_events.push_back(r.done);
    }
    d->ring_used.clear();
    if (d->ring) (void)hipHostFree(d->ring);
    d->ring = nullptr;
    d->ring_bytes = d->ring_head = 0;
}

// A region of `need` bytes of the staging ring at *off (false when the ring is
// absent or smaller than `need`). Regions whose uploads completed are freed
// first; while the ring is full, waits for the oldest.
static bool ring_take(slosh_data *d, size_t need, size_t *off) {
    need = (need + 255) & ~(size_t)255;
    if (!d->ring || need >= d->ring_bytes) return false;
    for (;;) {
        while (!d->ring_used.empty()) {
            const hipError_t q = hipEventQuery(d->ring_used.front().done);
            (void)hipGetLastError();  // a not-ready query must not linger as the last error
            if (q != hipSuccess) break;
            d->ring_events.push_back(d->ring_used.front().done);
            d->ring_used.pop_front();
        }
        if (d->ring_used.empty()) {
            *off = 0;
            return true;
        }
        // Used bytes run from the oldest region's start to ring_head (wrapping);
        // a new region must end strictly before the oldest one starts.
        const size_t tail = d->ring_used.front().off;
        if (d->ring_head > tail) {
            if (d->ring_bytes - d->ring_head >= need) {
                *off = d->ring_head;
                return true;
            }
            if (tail > need) {
                *off = 0;
                return true;
            }
        } else if (tail - d->ring_head > need) {
            *off = d->ring_head;
            return true;
        }
        (void)hipEventSynchronize(d->ring_used.front().done);
    }
}

// Marks [off, off + need) of the ring as in use until the work queued on `st`
// so far has completed.
static int ring_commit(slosh_data *d, size_t off, size_t need, hipStream_t st) {
    need = (need + 255) & ~(size_t)255;
    hipEvent_t ev = nullptr;
    if (!d->ring_events.empty()) {
        ev = d->ring_events.back();
        d->ring_events.pop_back();
    } else {
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(ev, st));
    d->ring_used.push_back({off, need, ev});
    d->ring_head = off + need;
    return SLOSH_OK;
}

static void free_data(slosh_data *d) {
    invalidate_graphs(d);
    for (slosh_graph *g : d->user_graphs) g->data = nullptr;  // their launches now fail cleanly
    d->user_graphs.clear();
    for (int i = 0; i < 4; i++)
        if (d->graph_done[i]) (void)hipEventDestroy(d->graph_done[i]);
    free_particle_buffers(d);
    for (int i = 0; i < 2; i++) {
        dfree(d->keys[i]);
        dfree(d->coll[i]);
        dfree(d->blk_start[i]);
        dfree(d->blk_cnt[i]);
        dfree(d->stay[i]);
        dfree(d->mcount[i]);
    }
    dfree(d->blist);
    dfree(d->nbr);
    dfree(d->cell_cnt);
    dfree(d->cell_start);
    dfree(d->mlist);
    dfree(d->bin);
    dfree(d->blk_tile_sums);
    dfree(d->counters);
    dfree(d->tiles);
    dfree(d->nodes);
    dfree(d->shapes);
    dfree(d->poses);
    dfree(d->vels);
    dfree(d->mprops);
    dfree(d->local_com);
    dfree(d->mats);
    dfree(d->impulses);
    dfree(d->dead);
    dfree(d->scene_rec);
    dfree(d->mesh_vtx);
    dfree(d->mesh_idx);
    dfree(d->dt_bound);
    release_ring(d);
    for (hipEvent_t e : d->ring_events) (void)hipEventDestroy(e);
    d->ring_events.clear();
    for (int i = 0; i < d->num_stage; i++) {
        if (d->stage[i].done) {
            (void)hipEventSynchronize(d->stage[i].done);
            (void)hipEventDestroy(d->stage[i].done);
        }
        if (d->stage[i].ptr) (void)hipHostFree(d->stage[i].ptr);
    }
    if (d->capture_stream) (void)hipStreamDestroy(d->capture_stream);
    delete d;
}

static bool body_may_move(const slosh_body_desc &b) {
    for (int i = 0; i < 3; i++)
        if (b.vel.linear[i] != 0.0f || b.vel.angular[i] != 0.0f || b.mprops.inv_mass[i] != 0.0f) return true;
    return false;
}

static int upload_bodies(slosh_data *d) {
    const uint32_t nb = d->nbodies;
    std::vector<float> shapes(8 * nb), poses(8 * nb), vels(8 * nb), mprops(20 * nb), lcom(4 * nb);
    std::vector<uint32_t> mats(2 * nb);
    for (uint32_t i = 0; i < nb; i++) {
        const slosh_body_desc &b = d->bodies_host[i];
        std::memcpy(&shapes[8 * i], &b.shape, 32);
        std::memcpy(&poses[8 * i], &b.pose, 32);
        for (int c = 0; c < 3; c++) {
            vels[8 * i + c] = b.vel.linear[c];
            vels[8 * i + 4 + c] = b.vel.angular[c];
            lcom[4 * i + c] = b.local_mprops.com[c];
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) mprops[20 * i + 4 * r + c] = b.mprops.inv_inertia[r][c];
        for (int c = 0; c < 3; c++) {
            mprops[20 * i + 12 + c] = b.mprops.inv_mass[c];
            mprops[20 * i + 16 + c] = b.mprops.com[c];
        }
    }
    (void)mats;
    if (nb == 0) return SLOSH_OK;
    HIP_TRY(hipMemcpy(d->shapes, shapes.data(), 32 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->poses, poses.data(), 32 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->vels, vels.data(), 32 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->mprops, mprops.data(), 80 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->local_com, lcom.data(), 16 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d->impulses, 0, sizeof(int) * 6 * nb));
    return SLOSH_OK;
}

// Every particle's primary block, and its "+1" neighbours, must have exact keys
// (block_associated_to_point, grid.slang:346-355, in the kernels' arithmetic).
// The block of particle i (a coordinate of 1 << 24 for a non-finite or huge one).
static void host_block_of(const slosh_particles *p, size_t i, float h, bool pow2, float inv, int b[3]) {
    for (int c = 0; c < 3; c++) {
        const float x = p->positions[4 * i + c];
        const float q = pow2 ? x * inv : x / h;
        const float a = std::nearbyint(q) - 1.0f;  // round half to even (default rounding mode)
        const float f = std::floor(a / 4.0f);
        b[c] = std::isfinite(f) && std::fabs(f) < 1.0e6f ? (int)f : (1 << 24);
    }
}

// The first particle of [lo, hi) outside the block-key range, or hi.
static size_t first_out_of_range(const slosh_particles *p, size_t lo, size_t hi, float h, uint32_t batched) {
    const bool pow2 = (f2u(h) & 0x807fffffu) == 0u && f2u(h) - 0x00800000u < 0x7e000000u;
    const float inv = 1.0f / h;
    for (size_t i = lo; i < hi; i++) {
        int b[3];
        host_block_of(p, i, h, pow2, inv, b);
        if (!block_in_key_range(batched, b[0], b[1], b[2])) return i;
    }
    return hi;
}

static int check_key_range(const slosh_particles *p, float h, uint32_t batched) {
    const size_t n = p->len, T = host_threads(n);
    std::vector<size_t> bad(T, n);
    std::vector<std::thread> workers;
    for (size_t t = 1; t < T; t++) {
        try {
            workers.emplace_back([&, t]() { bad[t] = first_out_of_range(p, n * t / T, n * (t + 1) / T, h, batched); });
        } catch (...) {
            bad[t] = first_out_of_range(p, n * t / T, n * (t + 1) / T, h, batched);
        }
    }
    bad[0] = first_out_of_range(p, 0, n / T, h, batched);
    for (auto &w : workers) w.join();
    size_t i = n;
    for (size_t t = 0; t < T && i == n; t++)
        if (bad[t] < n * (t + 1) / T) i = bad[t];
    if (i == n) return SLOSH_OK;
    int b[3];
    const bool pow2 = (f2u(h) & 0x807fffffu) == 0u && f2u(h) - 0x00800000u < 0x7e000000u;
    host_block_of(p, i, h, pow2, 1.0f / h, b);
    return fail(SLOSH_E_KEY_RANGE, "particle " + std::to_string(i) + " lies in block (" + std::to_string(b[0]) + ", " +
                                       std::to_string(b[1]) + ", " + std::to_string(b[2]) +
                                       ") outside the block-key range" + (batched ? " (batched: [-16, 46] per axis)" : ""));
}

static int validate_particles(const slosh_particles *p) {
    if (!p) return fail(SLOSH_E_INVALID, "null particles");
    if (p->len > 0 && (!p->positions || !p->kinematics || !p->def_grad || !p->properties || !p->models))
        return fail(SLOSH_E_INVALID, "null particle buffer");
    if (p->len >= 0xffffffffull / 2) return fail(SLOSH_E_INVALID, "too many particles");
    return SLOSH_OK;
}

static int create_impl(slosh_pipeline *pipeline, const slosh_sim_params *params, const slosh_particles *particles,
                       size_t particle_capacity, const slosh_body_desc *bodies, size_t num_bodies,
                       const slosh_boundary_condition *body_materials, float cell_width, uint32_t grid_capacity,
                       slosh_data *d) {
    TRY(validate_particles(particles));
    if (!params || !(cell_width > 0.0f)) return fail(SLOSH_E_INVALID, "invalid params or cell width");
    if (num_bodies > SLOSH_MAX_COLLIDERS) return fail(SLOSH_E_INVALID, "CPIC only supports up to 16 colliders");
    if (num_bodies > 0 && (!bodies || !body_materials)) return fail(SLOSH_E_INVALID, "null body buffers");
    if (grid_capacity == 0 || grid_capacity > (1u << 26)) return fail(SLOSH_E_INVALID, "grid capacity out of range");
    for (size_t i = 0; i < num_bodies; i++) {
        uint32_t ty = f2u(bodies[i].shape.a[3]);
        if (ty > 6u) return fail(SLOSH_E_UNSUPPORTED, "unknown collider shape type");
    }
    d->pipeline = pipeline;
    d->params = *params;
    d->h = cell_width;
    d->cap = next_pow2(grid_capacity);
    d->n = particles->len;
    d->pcap = std::max(particle_capacity, particles->len);
    if (d->pcap == 0) d->pcap = 1;
    d->batched = particles->scene_ids != nullptr;
    TRY(check_key_range(particles, cell_width, d->batched));
    bool any_force = false, any_sand = false;
    Staging s;
    TRY(stage_particles(particles, 0, s, &any_force, &any_sand));
    // Batched data keeps Drucker-Prager state arrays from the start, so scenes of
    // any material can be appended without reallocating.
    d->has_sand = (any_sand || d->batched) ? 1u : 0u;
    d->has_force = any_force ? 1u : 0u;
    TRY(alloc_particle_buffers(d, d->pcap, d->has_sand != 0));
    if (d->batched) TRY(dalloc(&d->scene_rec, 4 * (size_t)SLOSH_MAX_SCENES));
    note_append(d, particles);
    HIP_TRY(hipStreamCreateWithFlags(&d->capture_stream, hipStreamNonBlocking));
    TRY(upload_staging(d, s, 0, 0, 0, d->capture_stream));
    const size_t cap = d->cap;
    for (int i = 0; i < 2; i++) {
        TRY(dalloc(&d->keys[i], cap));
        HIP_TRY(hipMemset(d->keys[i], 0xff, 4 * cap));
        TRY(dalloc(&d->blk_start[i], cap));
        TRY(dalloc(&d->blk_cnt[i], cap));
        TRY(dalloc(&d->stay[i], cap));
        TRY(dalloc(&d->mcount[i], cap));
    }
    TRY(dalloc(&d->blist, cap));
    TRY(dalloc(&d->nbr, cap * NBR_STRIDE));
    TRY(dalloc(&d->cell_cnt, cap * NODES_PER_BLOCK));
    TRY(dalloc(&d->cell_start, cap * NODES_PER_BLOCK));
    TRY(dalloc(&d->mlist, cap * MOVER_CAP));
    TRY(dalloc(&d->bin, cap));
    TRY(dalloc(&d->blk_tile_sums, (cap + 4095) / 4096 + 1));
    TRY(dalloc(&d->counters, 16));
    HIP_TRY(hipMemset(d->counters, 0, 64));
    TRY(dalloc(&d->dead, SLOSH_MAX_SCENES + 1));
    TRY(dalloc(&d->dt_bound, 1));
    TRY(dalloc(&d->tiles, cap * TILE_NODES));
    TRY(dalloc(&d->nodes, cap * NODES_PER_BLOCK));
    d->nbodies = (uint32_t)num_bodies;
    for (size_t i = 0; i < num_bodies; i++) d->has_trimesh = d->has_trimesh || f2u(bodies[i].shape.a[3]) == 6u;
    d->bodies_host.assign(bodies, bodies + num_bodies);
    d->mats_host.assign(body_materials, body_materials + num_bodies);
    bool moving = false;
    for (size_t i = 0; i < num_bodies; i++) moving = moving || body_may_move(bodies[i]);
    d->coll_cache = moving ? 1u : 0u;
    d->coll_prev_valid = 1;  // the first substep's previous table is empty (grid.rs:454-457)
    if (d->coll_cache)
        for (int i = 0; i < 2; i++) TRY(dalloc(&d->coll[i], 2 * cap * NODES_PER_BLOCK));
    TRY(dalloc(&d->shapes, 2 * std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->poses, 2 * std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->vels, 2 * std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->mprops, 5 * std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->local_com, std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->mats, std::max<size_t>(num_bodies, 1)));
    TRY(dalloc(&d->impulses, 6 * std::max<size_t>(num_bodies, 1)));
    if (num_bodies) {
        std::vector<uint2> m(num_bodies);
        for (size_t i = 0; i < num_bodies; i++) m[i] = make_uint2(body_materials[i].ty, f2u(body_materials[i].friction));
        HIP_TRY(hipMemcpy(d->mats, m.data(), 8 * num_bodies, hipMemcpyHostToDevice));
    }
    TRY(upload_bodies(d));
    HIP_TRY(hipDeviceSynchronize());
    return SLOSH_OK;
}

// ---- Scratch of slosh_timestep_bound_noalloc ------------------------------
// One partial minimum per TSB_CHUNK particles: NOALLOC_PARTIALS words (64 KiB)
// on every GPU the library runs on, held until the process ends, shared by all
// calls on that GPU. In host code `scratch` is a host-side stand-in, not a
// device address: kernels get the address hipGetSymbolAddress returns for the
// device (noalloc_scratch[device]), resolved once per device at slosh_data_create
// with that device current.
#define NOALLOC_PARTIALS 16384u
static_assert((size_t)NOALLOC_PARTIALS * TSB_CHUNK == SLOSH_NOALLOC_MAX_PARTICLES, "scratch size");
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) uint32_t scratch[NOALLOC_PARTIALS];
static std::mutex noalloc_mu;
static uint32_t *noalloc_scratch[NOALLOC_MAX_DEVICES] = {};

static int resolve_noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(SLOSH_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_scratch[device]) return SLOSH_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(SLOSH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    noalloc_scratch[device] = (uint32_t *)ptr;
    return SLOSH_OK;
}

static uint32_t *noalloc_scratch_of(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    return noalloc_scratch[device];
}

int slosh_data_create(slosh_pipeline *pipeline, const slosh_sim_params *params, const slosh_particles *particles,
                      size_t particle_capacity, const slosh_body_desc *bodies, size_t num_bodies,
                      const slosh_boundary_condition *body_materials, float cell_width, uint32_t grid_capacity,
                      slosh_data **out) {
    if (!pipeline || !out) return fail(SLOSH_E_INVALID, "null pipeline or output");
    *out = nullptr;
    slosh_data *d = new slosh_data();
    int rc = create_impl(pipeline, params, particles, particle_capacity, bodies, num_bodies, body_materials,
                         cell_width, grid_capacity, d);
    if (rc == SLOSH_OK)
        rc = hipGetDevice(&d->device) == hipSuccess ? resolve_noalloc_scratch(d->device)
                                                    : fail(SLOSH_E_HIP, "hipGetDevice failed");
    if (rc != SLOSH_OK) {
        std::string msg = g_last_error;
        free_data(d);
        g_last_error = msg;
        return rc;
    }
    *out = d;
    return SLOSH_OK;
}

void slosh_data_destroy(slosh_data *d) {
    if (d) free_data(d);
}

int slosh_launch_step(slosh_pipeline *pipeline, slosh_data *d, slosh_stream stream, const slosh_hooks *hooks,
                      void *hooks_state, slosh_timestamps *ts) {
    if (!pipeline || !d) return fail(SLOSH_E_INVALID, "null pipeline or data");
    return run_step(pipeline, d, (hipStream_t)stream, hooks, hooks_state, ts);
}

int slosh_launch_steps_graph(slosh_pipeline *pipeline, slosh_data *d, slosh_stream stream, uint32_t count) {
    if (!pipeline || !d) return fail(SLOSH_E_INVALID, "null pipeline or data");
    hipStream_t st = (hipStream_t)stream;
    bool launched[4] = {false, false, false, false};
    auto mark = [&]() {
        for (int i = 0; i < 4; i++)
            if (launched[i]) mark_launch(&d->graph_done[i], st);
    };
    while (count >= 2) {
        // The substep alternates particle buffers and hash tables, so a pair of
        // substeps returns to the same parity; one graph per starting parity.
        const int slot = (d->cur_p & 1) | ((d->cur_g & 1) << 1);
        if (d->has_force || (d->coll_cache && !d->coll_prev_valid) || !d->next_primed) {
            const int rc = run_step(pipeline, d, st, nullptr, nullptr, nullptr);
            if (rc != SLOSH_OK) {
                mark();
                return rc;
            }
            count--;
            continue;
        }
        // A pair captured under other step parameters is stale (a guard behind
        // invalidate_graphs: every parameter a replay would reuse is compared).
        const StepParams now = make_params(d);
        if (d->graph[slot] && std::memcmp(&now, &d->graph_params[slot], sizeof now) != 0) {
            retire_exec(d->graph[slot], d->graph_done[slot]);
            launched[slot] = false;
        }
        if (!d->graph[slot]) {
            const int save_p = d->cur_p, save_g = d->cur_g;
            const uint32_t save_force = d->has_force;
            const bool save_primed = d->next_primed;
            HIP_TRY(hipStreamBeginCapture(d->capture_stream, hipStreamCaptureModeThreadLocal));
            int rc = run_step(pipeline, d, d->capture_stream, nullptr, nullptr, nullptr);
            if (rc == SLOSH_OK) rc = run_step(pipeline, d, d->capture_stream, nullptr, nullptr, nullptr);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(d->capture_stream, &graph);
            d->cur_p = save_p;
            d->cur_g = save_g;
            d->has_force = save_force;
            d->next_primed = save_primed;
            if (rc != SLOSH_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return fail(SLOSH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&d->graph[slot], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                d->graph[slot] = nullptr;
                return fail(SLOSH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            d->graph_params[slot] = now;
        }
        const hipError_t le = hipGraphLaunch(d->graph[slot], st);
        if (le != hipSuccess) {
            mark();
            return fail(SLOSH_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(le));
        }
        launched[slot] = true;
        count -= 2;
    }
    mark();
    if (count == 1) TRY(run_step(pipeline, d, st, nullptr, nullptr, nullptr));
    return SLOSH_OK;
}

int slosh_graph_capture(slosh_pipeline *pipeline, slosh_data *d, uint32_t steps, slosh_graph **out) {
    if (!pipeline || !d || !out) return fail(SLOSH_E_INVALID, "null argument");
    if (steps == 0 || (steps & 1u)) return fail(SLOSH_E_INVALID, "steps must be even and non-zero");
    if (d->has_force || (d->coll_cache && !d->coll_prev_valid) || !d->next_primed)
        return fail(SLOSH_E_INVALID, "run one slosh_launch_step before capturing (state not steady)");
    *out = nullptr;
    slosh_graph *g = new slosh_graph();
    g->params = make_params(d);
    const int save_p = d->cur_p, save_g = d->cur_g;
    const uint32_t save_force = d->has_force;
    const bool save_primed = d->next_primed;
    hipError_t e = hipStreamBeginCapture(d->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return fail(SLOSH_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    int rc = SLOSH_OK;
    for (uint32_t i = 0; i < steps && rc == SLOSH_OK; i++)
        rc = run_step(pipeline, d, d->capture_stream, nullptr, nullptr, nullptr);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(d->capture_stream, &graph);
    const bool back = d->cur_p == save_p && d->cur_g == save_g && d->next_primed == save_primed;
    d->cur_p = save_p;
    d->cur_g = save_g;
    d->has_force = save_force;
    d->next_primed = save_primed;
    if (rc == SLOSH_OK && e != hipSuccess) rc = fail(SLOSH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (rc == SLOSH_OK && !back) rc = fail(SLOSH_E_INVALID, "captured substeps do not return to the captured state");
    if (rc == SLOSH_OK) {
        e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(SLOSH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != SLOSH_OK) {
        slosh_graph_destroy(g);
        return rc;
    }
    g->data = d;
    g->generation = d->generation;
    g->cur_p = save_p;
    g->cur_g = save_g;
    d->user_graphs.push_back(g);
    *out = g;
    return SLOSH_OK;
}

int slosh_graph_launch(slosh_graph *g, slosh_stream stream) {
    if (!g) return fail(SLOSH_E_INVALID, "null graph");
    slosh_data *d = g->data;
    if (!d) return fail(SLOSH_E_INVALID, "the graph's data was destroyed");
    const StepParams now = make_params(d);
    if (!g->exec || d->generation != g->generation || d->cur_p != g->cur_p || d->cur_g != g->cur_g ||
        !d->next_primed || std::memcmp(&now, &g->params, sizeof now) != 0)
        return fail(SLOSH_E_INVALID, "graph no longer matches the data state (mutated or stepped an odd count)");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    mark_launch(&g->done, (hipStream_t)stream);
    return SLOSH_OK;
}

void slosh_graph_destroy(slosh_graph *g) {
    if (!g) return;
    if (g->data) {
        auto &v = g->data->user_graphs;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
    }
    retire_exec(g->exec, g->done);
    if (g->done) (void)hipEventDestroy(g->done);
    delete g;
}

int slosh_data_set_collision_mesh(slosh_data *d, const float *vertices4, size_t num_vertices,
                                  const uint32_t *indices, size_t num_indices) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    if ((num_vertices && !vertices4) || (num_indices && !indices)) return fail(SLOSH_E_INVALID, "null mesh buffers");
    constexpr uint32_t LEAF = 0xffffffffu;
    for (uint32_t b = 0; b < d->nbodies; b++) {
        const slosh_shape &sh = d->bodies_host[b].shape;
        if (f2u(sh.a[3]) != 6u) continue;
        const uint64_t vroot = f2u(sh.a[0]), iroot = f2u(sh.a[1]), nodes = f2u(sh.a[2]);
        const uint64_t ntri = f2u(sh.b[0]), nv = f2u(sh.b[1]);
        if (nodes == 0 || ntri == 0 || nv < 3 || nodes > (1u << 28) || ntri > (1u << 28) || nv > (1u << 28))
            return fail(SLOSH_E_INVALID, "trimesh body " + std::to_string(b) + ": empty or oversized mesh");
        if (vroot + 2 * nodes + 2 * nv + 3 * ntri > num_vertices || iroot + 3 * nodes + 3 * ntri > num_indices)
            return fail(SLOSH_E_INVALID, "trimesh body " + std::to_string(b) + ": ranges exceed the mesh buffers");
        for (uint64_t k = 0; k < nodes; k++) {
            const uint32_t *nd = indices + iroot + 3 * k;
            const bool ok_entry = nd[0] == LEAF ? nd[2] < ntri : (nd[0] > k && nd[0] < nodes);
            const bool ok_exit = nd[1] > k && nd[1] <= nodes;
            if (!ok_entry || !ok_exit)
                return fail(SLOSH_E_INVALID, "trimesh body " + std::to_string(b) + ": BVH node " + std::to_string(k) +
                                                 " links backward or out of range");
        }
        // The grid update culls the collider by the first (root) node's box, so it
        // must hold every triangle.
        const float *root_mn = vertices4 + 4 * vroot, *root_mx = vertices4 + 4 * (vroot + 1);
        for (uint64_t j = 0; j < 3 * ntri; j++) {
            const uint32_t v = indices[iroot + 3 * nodes + j];
            if (v >= nv)
                return fail(SLOSH_E_INVALID, "trimesh body " + std::to_string(b) + ": triangle vertex out of range");
            const float *p = vertices4 + 4 * (vroot + 2 * nodes + v);
            for (int c = 0; c < 3; c++)
                if (!(p[c] >= root_mn[c] && p[c] <= root_mx[c]))
                    return fail(SLOSH_E_INVALID, "trimesh body " + std::to_string(b) +
                                                     ": the root BVH box does not hold every triangle");
        }
    }
    dfree(d->mesh_vtx);
    dfree(d->mesh_idx);
    TRY(dalloc(&d->mesh_vtx, 4 * std::max<size_t>(num_vertices, 1)));
    TRY(dalloc(&d->mesh_idx, std::max<size_t>(num_indices, 1)));
    if (num_vertices) HIP_TRY(hipMemcpy(d->mesh_vtx, vertices4, 16 * num_vertices, hipMemcpyHostToDevice));
    if (num_indices) HIP_TRY(hipMemcpy(d->mesh_idx, indices, 4 * num_indices, hipMemcpyHostToDevice));
    d->mesh_nv = num_vertices;
    d->mesh_ni = num_indices;
    invalidate_graphs(d);
    return SLOSH_OK;
}

int slosh_set_svd_fp64(slosh_data *d, int enable) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    d->svd_f64 = enable ? 1u : 0u;
    invalidate_graphs(d);
    return SLOSH_OK;
}

int slosh_set_sim_params(slosh_data *d, const slosh_sim_params *params) {
    if (!d || !params) return fail(SLOSH_E_INVALID, "null argument");
    d->params = *params;
    invalidate_graphs(d);
    return SLOSH_OK;
}

int slosh_timestep_bound(slosh_data *d, slosh_stream stream, float *out_dt, uint32_t *out_raw) {
    if (!d || !out_dt) return fail(SLOSH_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemsetD32Async(d->dt_bound, (int)0xffffffffu, 1, st));  // reset_timestep_bound
    const StepParams P = make_params(d);
    HIP_TRY(launch_timestep_bound(P, pbufs(d, d->cur_p), sbufs(d), (uint32_t)d->pcap, d->dt_bound, st));
    uint32_t raw = 0xffffffffu;
    HIP_TRY(hipMemcpyAsync(&raw, d->dt_bound, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *out_dt = (float)raw / 1.0e12f;  // GpuTimestepBounds::computed_dt
    if (out_raw) *out_raw = raw;
    return SLOSH_OK;
}

int slosh_timestep_bound_noalloc(slosh_data *d, slosh_stream stream, uint32_t *out_raw) {
    if (!d || !out_raw) return fail(SLOSH_E_INVALID, "null argument");
    if (d->n > (size_t)NOALLOC_PARTIALS * TSB_CHUNK)
        return fail(SLOSH_E_INVALID, "slosh_timestep_bound_noalloc: " + std::to_string(d->n) +
                                         " particles exceed the " +
                                         std::to_string((size_t)NOALLOC_PARTIALS * TSB_CHUNK) +
                                         " its static scratch holds (use slosh_timestep_bound)");
    uint32_t *ptr = noalloc_scratch_of(d->device);
    if (!ptr) return fail(SLOSH_E_HIP, "slosh_timestep_bound_noalloc: scratch address not resolved for the device");
    const StepParams P = make_params(d);
    HIP_TRY(launch_timestep_bound_partials(P, pbufs(d, d->cur_p), sbufs(d), (uint32_t)d->pcap, ptr, out_raw,
                                           (hipStream_t)stream));
    return SLOSH_OK;
}

int slosh_adaptive_substeps(slosh_data *d, slosh_stream stream, float base_dt, uint32_t min_substeps,
                            uint32_t max_substeps, uint32_t *out_substeps) {
    if (!d || !out_substeps) return fail(SLOSH_E_INVALID, "null argument");
    if (max_substeps == 0 || min_substeps > max_substeps) return fail(SLOSH_E_INVALID, "invalid substep range");
    uint32_t n = max_substeps;
    if (min_substeps < max_substeps) {
        float bound = 0.0f;
        TRY(slosh_timestep_bound(d, stream, &bound, nullptr));
        const float q = ceilf(base_dt / bound);
        // `as u32`: NaN -> 0, saturating at both ends.
        const uint32_t est = !(q > 0.0f) ? 0u : (q >= 4294967295.0f ? 0xffffffffu : (uint32_t)q);
        n = std::min(std::max(est, min_substeps), max_substeps);
    }
    const float dt = base_dt / (float)n;
    if (dt != d->params.dt) {
        d->params.dt = dt;
        invalidate_graphs(d);
    }
    *out_substeps = n;
    return SLOSH_OK;
}

int slosh_check_errors(slosh_data *d, slosh_stream stream) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    hipStream_t st = (hipStream_t)stream;
    uint32_t c[16];
    HIP_TRY(hipMemcpyAsync(c, d->counters, 64, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t err = c[CTR_ERRORS] | c[8 + CTR_ERRORS];
    if (err != 0u) {
        uint32_t zero = 0;
        HIP_TRY(hipMemcpy(d->counters + CTR_ERRORS, &zero, 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d->counters + 8 + CTR_ERRORS, &zero, 4, hipMemcpyHostToDevice));
        if (err & ERR_KEY_RANGE)
            return fail(SLOSH_E_KEY_RANGE, d->batched ? "a particle left its scene's block-key range (batched "
                                                        "blocks must stay in [-16, 46] per axis)"
                                                      : "a particle left the block-key range");
        return fail(SLOSH_E_GRID_OVERFLOW, "sparse-grid hash map overflow (increase grid_capacity)");
    }
    return SLOSH_OK;
}

size_t slosh_num_particles(const slosh_data *d) { return d ? d->n : 0; }

int slosh_read_positions(slosh_data *d, slosh_stream stream, float *out4) {
    if (!d || (!out4 && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    Snapshot s;
    TRY(download_current(d, (hipStream_t)stream, s));
    for (size_t i = 0; i < d->n; i++) {
        uint32_t id = s.ids[i];
        out4[4 * id + 0] = s.pos[4 * i + 0];
        out4[4 * id + 1] = s.pos[4 * i + 1];
        out4[4 * id + 2] = s.pos[4 * i + 2];
        out4[4 * id + 3] = 0.0f;
    }
    return SLOSH_OK;
}

int slosh_read_kinematics(slosh_data *d, slosh_stream stream, slosh_kinematics *out) {
    if (!d || (!out && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    Snapshot s;
    TRY(download_current(d, (hipStream_t)stream, s));
    const size_t n = d->n;
    for (size_t i = 0; i < n; i++) {
        slosh_kinematics &k = out[s.ids[i]];
        std::memset(&k, 0, sizeof k);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) k.affine.m[r][c] = s.C[(size_t)(r * 3 + c) * n + i];
        for (int c = 0; c < 3; c++) {
            k.velocity[c] = s.vm[4 * i + c];
            k.force_dt[c] = s.fdt[(size_t)c * n + i];
        }
        k.vel_grad_det = s.vgd[i];
        k.mass = s.vm[4 * i + 3];
        k.enabled = f2u(s.pos[4 * i + 3]) & 1u;
    }
    return SLOSH_OK;
}

int slosh_read_def_grad(slosh_data *d, slosh_stream stream, slosh_matrix *out) {
    if (!d || (!out && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    Snapshot s;
    TRY(download_current(d, (hipStream_t)stream, s));
    const size_t n = d->n;
    for (size_t i = 0; i < n; i++) {
        slosh_matrix &m = out[s.ids[i]];
        std::memset(&m, 0, sizeof m);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) m.m[r][c] = s.F[(size_t)(r * 3 + c) * n + i];
    }
    return SLOSH_OK;
}

int slosh_read_models(slosh_data *d, slosh_stream stream, slosh_particle_model *out) {
    if (!d || (!out && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    Snapshot s;
    TRY(download_current(d, (hipStream_t)stream, s));
    const size_t n = d->n;
    const std::vector<slosh_particle_model> models = flat_models(d);
    for (size_t i = 0; i < n; i++) {
        uint32_t id = s.ids[i];
        out[id] = models[id];
        if (model_is_sand(out[id]))
            for (int c = 0; c < 3; c++) out[id].words[1 + c] = f2u(s.dps[(size_t)c * n + i]);
    }
    return SLOSH_OK;
}

int slosh_read_properties(slosh_data *d, slosh_stream stream, slosh_particle_properties *out) {
    (void)stream;
    if (!d || (!out && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    size_t o = 0;
    for (const auto &r : d->runs) {
        if (r.uniform) std::fill(out + o, out + o + r.count, r.recs_p[0]);
        else std::copy(r.recs_p.begin(), r.recs_p.end(), out + o);
        o += r.count;
    }
    return SLOSH_OK;
}

int slosh_read_scene_ids(slosh_data *d, slosh_stream stream, uint32_t *out) {
    (void)stream;
    if (!d || (!out && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    size_t o = 0;
    for (const auto &r : d->runs) {
        std::fill(out + o, out + o + r.count, r.scene);
        o += r.count;
    }
    return SLOSH_OK;
}

int slosh_read_bodies(slosh_data *d, slosh_stream stream, slosh_sim3 *poses, slosh_velocity *vels) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    const uint32_t nb = d->nbodies;
    if (nb == 0) return SLOSH_OK;
    hipStream_t st = (hipStream_t)stream;
    std::vector<float> p(8 * nb), v(8 * nb);
    HIP_TRY(hipMemcpyAsync(p.data(), d->poses, 32 * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(v.data(), d->vels, 32 * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < nb; i++) {
        if (poses) std::memcpy(&poses[i], &p[8 * i], 32);
        if (vels) {
            std::memset(&vels[i], 0, sizeof(slosh_velocity));
            for (int c = 0; c < 3; c++) {
                vels[i].linear[c] = v[8 * i + c];
                vels[i].angular[c] = v[8 * i + 4 + c];
            }
        }
    }
    return SLOSH_OK;
}

int slosh_num_blocks(slosh_data *d, slosh_stream stream, uint32_t *out) {
    if (!d || !out) return fail(SLOSH_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    uint32_t c[8];
    HIP_TRY(hipMemcpyAsync(c, ctr_set(d, d->cur_g), 32, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *out = c[CTR_NUM_BLOCKS];
    return SLOSH_OK;
}

static const int ASSOC_H[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

// The current table's block list (header ids = table slots), first B entries.
static int read_block_list(slosh_data *d, hipStream_t st, uint32_t B, std::vector<uint32_t> &hids) {
    hids.resize(B);
    if (B) {
        HIP_TRY(hipMemcpyAsync(hids.data(), d->blist, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return SLOSH_OK;
}

int slosh_read_blocks(slosh_data *d, slosh_stream stream, slosh_block_info *out, uint32_t max_blocks,
                      uint32_t *num_written) {
    if (!d || !out) return fail(SLOSH_E_INVALID, "null argument");
    uint32_t B = 0;
    TRY(slosh_num_blocks(d, stream, &B));
    hipStream_t st = (hipStream_t)stream;
    const size_t cap = d->cap;
    std::vector<uint32_t> hids;
    TRY(read_block_list(d, st, B, hids));  // every block: the extras need the neighbours' counts
    std::vector<uint32_t> keys(cap), cnt(cap * 64), nbr(cap * NBR_STRIDE), start(cap);
    if (B) {
        HIP_TRY(hipMemcpyAsync(keys.data(), d->keys[d->cur_g], 4 * cap, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(cnt.data(), d->cell_cnt, 4 * cap * 64, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(nbr.data(), d->nbr, 4 * cap * NBR_STRIDE, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(start.data(), d->blk_start[d->cur_g], 4 * cap, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    B = std::min(B, max_blocks);
    for (uint32_t b = 0; b < B; b++) {
        const uint32_t hid = hids[b];
        slosh_block_info &bi = out[b];
        std::memset(&bi, 0, sizeof bi);
        uint32_t scene;
        unpack_key(d->batched, keys[hid], &bi.vid[0], &bi.vid[1], &bi.vid[2], &scene);
        bi.first_particle = start[hid];
        // Primary buckets: the associated-cell slab along z (sort.slang:15-21).
        for (int c = 0; c < 64; c++) bi.bucket_counts[c >> 4] += cnt[(size_t)hid * 64 + c];
        // Extras received from the -1 neighbour B - o: its particles whose
        // associated cell is >= 2 along every axis where o = 1 (sort.slang:27-42).
        for (int o = 1; o < 8; o++) {
            uint32_t src = nbr[(size_t)hid * NBR_STRIDE + NBR_MINUS + o - 1];
            if (src == SLOSH_NONE || src >= cap) continue;
            for (int c = 0; c < 64; c++) {
                int lx = c & 3, ly = (c >> 2) & 3, lz = c >> 4;
                if ((ASSOC_H[o][0] && lx < 2) || (ASSOC_H[o][1] && ly < 2) || (ASSOC_H[o][2] && lz < 2)) continue;
                int local = lz - ASSOC_H[o][2] * 4;
                int bucket = 4 + std::max(local, -2) + 2;
                bi.bucket_counts[bucket] += cnt[(size_t)src * 64 + c];
            }
        }
        for (int k = 0; k < 4; k++) bi.num_particles += bi.bucket_counts[k];
        bi.num_particles_with_extras = bi.num_particles;
        for (int k = 4; k < 10; k++) bi.num_particles_with_extras += bi.bucket_counts[k];
    }
    if (num_written) *num_written = B;
    return SLOSH_OK;
}

int slosh_read_nodes(slosh_data *d, slosh_stream stream, float *out4, uint32_t max_blocks) {
    if (!d || !out4) return fail(SLOSH_E_INVALID, "null argument");
    uint32_t B = 0;
    TRY(slosh_num_blocks(d, stream, &B));
    B = std::min(B, max_blocks);
    hipStream_t st = (hipStream_t)stream;
    std::vector<uint32_t> hids;
    TRY(read_block_list(d, st, B, hids));
    if (B == 0) return SLOSH_OK;
    std::vector<float> all((size_t)d->cap * NODES_PER_BLOCK * 4);
    HIP_TRY(hipMemcpyAsync(all.data(), d->nodes, all.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t b = 0; b < B; b++)
        std::memcpy(out4 + (size_t)b * 256, all.data() + (size_t)hids[b] * 256, 1024);
    return SLOSH_OK;
}

int slosh_read_block_scenes(slosh_data *d, slosh_stream stream, uint32_t *out, uint32_t max_blocks) {
    if (!d || !out) return fail(SLOSH_E_INVALID, "null argument");
    uint32_t B = 0;
    TRY(slosh_num_blocks(d, stream, &B));
    B = std::min(B, max_blocks);
    hipStream_t st = (hipStream_t)stream;
    std::vector<uint32_t> hids, keys(d->cap);
    TRY(read_block_list(d, st, B, hids));
    if (B) {
        HIP_TRY(hipMemcpyAsync(keys.data(), d->keys[d->cur_g], 4 * (size_t)d->cap, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    for (uint32_t b = 0; b < B; b++) {
        int x, y, z;
        unpack_key(d->batched, keys[hids[b]], &x, &y, &z, &out[b]);
    }
    return SLOSH_OK;
}

// Rebuilds every particle buffer from host arrays in original order (used by
// append beyond capacity and by removals).
static int rebuild(slosh_data *d, hipStream_t st, size_t new_cap, const Staging &all, bool any_sand) {
    invalidate_graphs(d);
    HIP_TRY(hipStreamSynchronize(st));
    free_particle_buffers(d);
    d->pcap = std::max<size_t>(new_cap, std::max<size_t>(all.n, 1));
    d->n = all.n;
    d->has_sand = (d->has_sand || any_sand) ? 1u : 0u;
    TRY(alloc_particle_buffers(d, d->pcap, d->has_sand != 0));
    d->cur_p = 0;
    return upload_staging(d, all, 0, 0, 0, st);
}

// Current state as a Staging in original order (static parts from host copies).
// Current state in the reference layouts, original order.
struct RefArrays {
    std::vector<slosh_kinematics> kin;
    std::vector<slosh_matrix> F;
    std::vector<float> pos;
    std::vector<slosh_particle_model> models;
    std::vector<slosh_particle_properties> props;
    std::vector<uint32_t> scenes;
    slosh_particles view(bool batched) const {
        slosh_particles p;
        p.len = kin.size();
        p.positions = pos.data();
        p.kinematics = kin.data();
        p.def_grad = F.data();
        p.properties = props.data();
        p.models = models.data();
        p.scene_ids = batched ? scenes.data() : nullptr;
        return p;
    }
};

static int gather_reference(slosh_data *d, hipStream_t st, RefArrays &r) {
    Snapshot s;
    TRY(download_current(d, st, s));
    const size_t n = d->n;
    r.kin.resize(n);
    r.F.resize(n);
    r.pos.assign(4 * n, 0.0f);
    r.models.resize(n);
    const std::vector<slosh_particle_model> old_models = flat_models(d);
    for (size_t i = 0; i < n; i++) {
        uint32_t id = s.ids[i];
        for (int c = 0; c < 3; c++) r.pos[4 * id + c] = s.pos[4 * i + c];
        slosh_kinematics &k = r.kin[id];
        std::memset(&k, 0, sizeof k);
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
                k.affine.m[a][c] = s.C[(size_t)(a * 3 + c) * n + i];
                r.F[id].m[a][c] = s.F[(size_t)(a * 3 + c) * n + i];
            }
        for (int c = 0; c < 3; c++) {
            k.velocity[c] = s.vm[4 * i + c];
            k.force_dt[c] = s.fdt[(size_t)c * n + i];
        }
        k.vel_grad_det = s.vgd[i];
        k.mass = s.vm[4 * i + 3];
        k.enabled = f2u(s.pos[4 * i + 3]) & 1u;
        r.models[id] = old_models[id];
        if (model_is_sand(r.models[id]))
            for (int c = 0; c < 3; c++) r.models[id].words[1 + c] = f2u(s.dps[(size_t)c * n + i]);
    }
    set_models(d, r.models);
    r.props = flat_props(d);
    r.scenes = flat_scenes(d);
    return SLOSH_OK;
}

// Current state as a Staging in original order.
static int gather_all(slosh_data *d, hipStream_t st, Staging &all, bool *any_force) {
    RefArrays r;
    TRY(gather_reference(d, st, r));
    const slosh_particles p = r.view(d->batched != 0);
    bool any_sand = false;
    return stage_particles(&p, 0, all, any_force, &any_sand);
}

static void concat_staging(Staging &a, const Staging &b) {
    Staging r;
    r.resize(a.n + b.n);
    auto cat4 = [&](std::vector<float> &dst, const std::vector<float> &x, const std::vector<float> &y) {
        std::copy(x.begin(), x.end(), dst.begin());
        std::copy(y.begin(), y.end(), dst.begin() + x.size());
    };
    cat4(r.pos, a.pos, b.pos);
    cat4(r.vm, a.vm, b.vm);
    cat4(r.prop, a.prop, b.prop);
    cat4(r.model, a.model, b.model);
    cat4(r.dp0, a.dp0, b.dp0);
    cat4(r.dp1, a.dp1, b.dp1);
    cat4(r.vgd, a.vgd, b.vgd);
    auto catk = [&](std::vector<float> &dst, const std::vector<float> &x, const std::vector<float> &y, int comps) {
        for (int k = 0; k < comps; k++) {
            std::copy(x.begin() + k * a.n, x.begin() + (k + 1) * a.n, dst.begin() + k * r.n);
            std::copy(y.begin() + k * b.n, y.begin() + (k + 1) * b.n, dst.begin() + k * r.n + a.n);
        }
    };
    catk(r.C, a.C, b.C, 9);
    catk(r.F, a.F, b.F, 9);
    catk(r.dps, a.dps, b.dps, 3);
    catk(r.fdt, a.fdt, b.fdt, 3);
    std::copy(a.ids.begin(), a.ids.end(), r.ids.begin());
    for (size_t i = 0; i < b.n; i++) r.ids[a.n + i] = (uint32_t)(a.n + i);
    a = std::move(r);
}

// A pinned arena of at least `need` bytes whose earlier uploads are done: an idle
// one that fits, else an idle one grown, else a new one (up to MAX_STAGE_ARENAS),
// else the first, once its uploads complete. Pinning is slow, so arenas grow
// geometrically.
static int pick_stage_arena(slosh_data *d, size_t need, slosh_data::StageArena **out) {
    slosh_data::StageArena *fit = nullptr, *idle = nullptr;
    for (int i = 0; i < d->num_stage; i++) {
        slosh_data::StageArena &a = d->stage[i];
        bool ready = true;
        if (a.done) {
            ready = hipEventQuery(a.done) == hipSuccess;
            (void)hipGetLastError();  // a not-ready query must not linger as the last error
        }
        if (!ready) continue;
        if (a.bytes >= need && (!fit || a.bytes < fit->bytes)) fit = &a;
        if (!idle) idle = &a;
    }
    slosh_data::StageArena *a = fit ? fit : idle;
    if (!a && d->num_stage < slosh_data::MAX_STAGE_ARENAS) a = &d->stage[d->num_stage++];
    if (!a) {
        a = &d->stage[0];
        HIP_TRY(hipEventSynchronize(a->done));
    }
    if (a->bytes < need) {
        const size_t bytes = std::max(need, a->bytes + a->bytes / 2);
        if (a->ptr) HIP_TRY(hipHostFree(a->ptr));
        a->ptr = nullptr;
        a->bytes = 0;
        HIP_TRY(hipHostMalloc((void **)&a->ptr, bytes, hipHostMallocDefault));
        a->bytes = bytes;
    }
    if (!a->done) HIP_TRY(hipEventCreateWithFlags(&a->done, hipEventDisableTiming));
    *out = a;
    return SLOSH_OK;
}

int slosh_reserve_staging(slosh_data *d, size_t bytes) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    release_ring(d);
    if (bytes == 0) return SLOSH_OK;
    HIP_TRY(hipHostMalloc((void **)&d->ring, bytes, hipHostMallocDefault));
    d->ring_bytes = bytes;
    return SLOSH_OK;
}

int slosh_append_particles(slosh_data *d, slosh_stream stream, const slosh_particles *particles) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    TRY(validate_particles(particles));
    if ((particles->scene_ids != nullptr) != (d->batched != 0))
        return fail(SLOSH_E_INVALID, "scene ids must be given iff the data is batched");
    TRY(check_key_range(particles, d->h, d->batched));
    hipStream_t st = (hipStream_t)stream;
    if (particles->len == 0) return SLOSH_OK;
    bool any_force = false, any_sand = false;
    bool need_sand = false;
    for (size_t i = 0; i < particles->len && !need_sand && !d->has_sand; i++)
        need_sand = model_is_sand(particles->models[i]);
    if (d->n + particles->len <= d->pcap && !need_sand) {
        // Into the free tail: stage in a pinned arena and do the host bookkeeping
        // while substeps already queued on `stream` run, then queue the copies on
        // `stream` (the next substeps follow them) without waiting for them. Only
        // captured graphs need the queued substeps to finish first (they are
        // released); without graphs the host goes on at once.
        const size_t need = PinnedStaging::WORDS * 4 * particles->len;
        slosh_data::StageArena *arena = nullptr;
        size_t ring_off = 0;
        const bool in_ring = ring_take(d, need, &ring_off);
        if (!in_ring) TRY(pick_stage_arena(d, need, &arena));
        PinnedStaging s(in_ring ? d->ring + ring_off : arena->ptr, particles->len);
        TRY(stage_particles(particles, (uint32_t)d->n, s, &any_force, &any_sand));
        note_append(d, particles);
        if (d->graph[0] || d->graph[1] || d->graph[2] || d->graph[3]) HIP_TRY(hipStreamSynchronize(st));
        invalidate_graphs(d);
        TRY(upload_staging(d, s, d->cur_p, d->n, d->n, st, false));
        if (in_ring) TRY(ring_commit(d, ring_off, need, st));
        else HIP_TRY(hipEventRecord(arena->done, st));
        d->n += particles->len;
        if (any_force) d->has_force = 1;
        return SLOSH_OK;
    } else {
        Staging s;
        TRY(stage_particles(particles, (uint32_t)d->n, s, &any_force, &any_sand));
        HIP_TRY(hipStreamSynchronize(st));
        invalidate_graphs(d);
        // GpuTensor::append reallocates; we rebuild in original order with headroom
        // (rebuild sets d->n to the combined count).
        Staging all;
        bool f2 = false;
        TRY(gather_all(d, st, all, &f2));
        concat_staging(all, s);
        any_force = any_force || f2;
        TRY(rebuild(d, st, (d->n + particles->len) * 3 / 2 + 1024, all, any_sand));
    }
    note_append(d, particles);
    if (any_force) d->has_force = 1;
    return SLOSH_OK;
}

// Device-side removal (mode 0: the scenes flagged in dead_host; mode 1: ids in
// [lo, hi)). Kept particles are compacted into the other particle buffer in slot
// order and their original ids renumbered in order; the host runs follow.
static int device_remove(slosh_data *d, hipStream_t st, uint32_t mode, const std::vector<uint8_t> &dead_host,
                         size_t lo, size_t hi) {
    const size_t n = d->n;
    invalidate_graphs(d);
    if (mode == 0) {
        HIP_TRY(hipStreamSynchronize(st));  // a previous flag upload may still read dead_host
        d->dead_host = dead_host;
        HIP_TRY(hipMemcpyAsync(d->dead, d->dead_host.data(), SLOSH_MAX_SCENES, hipMemcpyHostToDevice, st));
    }
    const int c = d->cur_p, o = c ^ 1;
    CompactArgs a{};
    for (int k = 0; k < 2; k++) {
        const PBufs &p = d->pb[k == 0 ? c : o];
        a.pos[k] = p.pos;
        a.vm[k] = p.vm;
        a.C[k] = p.C;
        a.F[k] = p.F;
        a.vgd[k] = p.vgd;
        a.dps[k] = p.dps;
        a.ids[k] = p.ids;
    }
    a.fdt[0] = d->fdt;
    a.fdt[1] = d->fdt2;
    a.prop[0] = d->s_prop;
    a.prop[1] = d->s2_prop;
    a.model[0] = d->s_model;
    a.model[1] = d->s2_model;
    a.dp0[0] = d->s_dp0;
    a.dp0[1] = d->s2_dp0;
    a.dp1[0] = d->s_dp1;
    a.dp1[1] = d->s2_dp1;
    a.keep_slot = d->pslot;
    a.keep_id = d->prank;
    a.keep_id2 = d->perm;
    a.tile_sums = d->scan_tmp;
    a.total = d->scan_tmp + scan_tiles((uint32_t)d->pcap);
    HIP_TRY(launch_compact(a, (uint32_t)n, (uint32_t)d->pcap, mode, d->dead, (uint32_t)lo, (uint32_t)hi, st));
    d->cur_p = o;
    std::swap(d->fdt, d->fdt2);
    std::swap(d->s_prop, d->s2_prop);
    std::swap(d->s_model, d->s2_model);
    std::swap(d->s_dp0, d->s2_dp0);
    std::swap(d->s_dp1, d->s2_dp1);
    // Host runs: drop removed runs (mode 0) or cut [lo, hi) out of them (mode 1).
    std::vector<HostRun> kept;
    size_t off = 0, m = 0;
    for (auto &r : d->runs) {
        const size_t len = r.count;
        if (mode == 0) {
            if (dead_host[r.scene]) {
                unnote_run(d, r, 0, len);
                forget_scene(d, r.scene);
            } else {
                m += len;
                kept.push_back(std::move(r));
            }
        } else {
            const size_t a0 = std::max(lo, off), a1 = std::min(hi, off + len);
            if (a0 < a1) {
                unnote_run(d, r, a0 - off, a1 - off);
                if (!r.uniform) {
                    r.recs_m.erase(r.recs_m.begin() + (a0 - off), r.recs_m.begin() + (a1 - off));
                    r.recs_p.erase(r.recs_p.begin() + (a0 - off), r.recs_p.begin() + (a1 - off));
                }
                r.count -= a1 - a0;
            }
            if (r.count > 0) {
                m += r.count;
                if (!kept.empty() && kept.back().scene == r.scene) {
                    HostRun &k = kept.back();
                    if (!(k.uniform && r.uniform && same_records(k.recs_m[0], k.recs_p[0], r.recs_m[0], r.recs_p[0]))) {
                        k.expand();
                        r.expand();
                        k.recs_m.insert(k.recs_m.end(), r.recs_m.begin(), r.recs_m.end());
                        k.recs_p.insert(k.recs_p.end(), r.recs_p.begin(), r.recs_p.end());
                    }
                    k.count += r.count;
                } else {
                    kept.push_back(std::move(r));
                }
            }
        }
        off += len;
    }
    d->runs = std::move(kept);
    d->n = m;
    note_records(d, nullptr, nullptr, 0, -1);  // refresh the uniform state once the count is known
    return SLOSH_OK;
}

int slosh_remove_particles(slosh_data *d, slosh_stream stream, size_t start, size_t end) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    if (start > end || end > d->n) return fail(SLOSH_E_INVALID, "removal range out of bounds");
    if (start == end) return SLOSH_OK;
    return device_remove(d, (hipStream_t)stream, 1, {}, start, end);
}

int slosh_remove_scenes(slosh_data *d, slosh_stream stream, const uint32_t *scene_ids, size_t count,
                        size_t *removed) {
    if (!d || (!scene_ids && count)) return fail(SLOSH_E_INVALID, "null argument");
    std::vector<uint8_t> dead(SLOSH_MAX_SCENES, 0);
    for (size_t k = 0; k < count; k++) {
        if (scene_ids[k] >= SLOSH_MAX_SCENES) return fail(SLOSH_E_INVALID, "scene id out of range");
        dead[scene_ids[k]] = 1;
    }
    size_t cnt = 0;
    for (const auto &r : d->runs)
        if (dead[r.scene]) cnt += r.count;
    if (removed) *removed = cnt;
    if (cnt == 0) return SLOSH_OK;
    return device_remove(d, (hipStream_t)stream, 0, dead, 0, 0);
}

int slosh_remove_scene(slosh_data *d, slosh_stream stream, uint32_t scene_id, size_t *removed) {
    return slosh_remove_scenes(d, stream, &scene_id, 1, removed);
}

size_t slosh_scene_size(const slosh_data *d, uint32_t scene_id) {
    if (!d) return 0;
    size_t c = 0;
    for (const auto &r : d->runs)
        if (r.scene == scene_id) c += r.count;
    return c;
}

int slosh_read_scene(slosh_data *d, slosh_stream stream, uint32_t scene_id, float *pos4, slosh_kinematics *kin,
                     slosh_matrix *def_grad, size_t max_count, size_t *count) {
    if (!d) return fail(SLOSH_E_INVALID, "null data");
    const size_t m = slosh_scene_size(d, scene_id);
    if (count) *count = m;
    if (m == 0) return SLOSH_OK;
    if (m > max_count) return fail(SLOSH_E_INVALID, "output smaller than the scene");
    hipStream_t st = (hipStream_t)stream;
    // Staging: the particle buffer the next substep writes (free between substeps).
    const PBufs &cur = d->pb[d->cur_p];
    const PBufs &stg = d->pb[d->cur_p ^ 1];
    GatherArgs a{};
    a.pos = cur.pos;
    a.vm = cur.vm;
    a.C = cur.C;
    a.F = cur.F;
    a.vgd = cur.vgd;
    a.dps = cur.dps;
    a.fdt = d->fdt;
    a.ids = cur.ids;
    a.s_pos = stg.pos;
    a.s_vm = stg.vm;
    a.s_C = stg.C;
    a.s_F = stg.F;
    a.s_vgd = stg.vgd;
    a.s_dps = stg.dps;
    a.s_fdt = d->fdt2;
    a.rank_id = d->prank;
    a.tile_sums = d->scan_tmp;
    a.total = d->scan_tmp + scan_tiles((uint32_t)d->pcap);
    HIP_TRY(launch_scene_gather(a, (uint32_t)d->n, (uint32_t)d->pcap, scene_id, (uint32_t)m, st));
    // Positions go straight to the caller's buffer (their w, the flag bits, is
    // cleared below); the other fields through temporaries allocated only when
    // asked for and left uninitialised (the copies fill them).
    std::unique_ptr<float[]> pos_tmp, vm, C, F, vgd, fdt;
    float *pos = pos4;
    if (!pos && kin) {
        pos_tmp.reset(new float[4 * m]);
        pos = pos_tmp.get();
    }
    if (pos) HIP_TRY(hipMemcpyAsync(pos, stg.pos, 16 * m, hipMemcpyDeviceToHost, st));
    if (kin) {
        vm.reset(new float[4 * m]);
        C.reset(new float[9 * m]);
        vgd.reset(new float[m]);
        fdt.reset(new float[3 * m]);
        HIP_TRY(hipMemcpyAsync(vm.get(), stg.vm, 16 * m, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(C.get(), stg.C, 36 * m, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(vgd.get(), stg.vgd, 4 * m, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(fdt.get(), d->fdt2, 12 * m, hipMemcpyDeviceToHost, st));
    }
    if (def_grad) {
        F.reset(new float[9 * m]);
        HIP_TRY(hipMemcpyAsync(F.get(), stg.F, 36 * m, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t j = 0; j < m; j++) {
        if (kin) {
            slosh_kinematics &k = kin[j];
            std::memset(&k, 0, sizeof k);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) k.affine.m[r][c] = C[(size_t)(r * 3 + c) * m + j];
            for (int c = 0; c < 3; c++) {
                k.velocity[c] = vm[4 * j + c];
                k.force_dt[c] = fdt[(size_t)c * m + j];
            }
            k.vel_grad_det = vgd[j];
            k.mass = vm[4 * j + 3];
            k.enabled = f2u(pos[4 * j + 3]) & 1u;
        }
        if (pos4) pos4[4 * j + 3] = 0.0f;
        if (def_grad) {
            slosh_matrix &mm = def_grad[j];
            std::memset(&mm, 0, sizeof mm);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) mm.m[r][c] = F[(size_t)(r * 3 + c) * m + j];
        }
    }
    return SLOSH_OK;
}

// ---- launch-shape tuning ------------------------------------------------------
static bool tuning_valid(const slosh_tuning &t) {
    return (t.p2g_chunk == 192u || t.p2g_chunk == 256u || t.p2g_chunk == 384u || t.p2g_chunk == 512u) &&
           (t.g2p_threads == 128u || t.g2p_threads == 256u) && t.block_wgs >= 64u && t.block_wgs <= 65536u;
}

int slosh_set_tuning(slosh_data *d, const slosh_tuning *t) {
    if (!d || !t) return fail(SLOSH_E_INVALID, "null argument");
    if (!tuning_valid(*t)) return fail(SLOSH_E_INVALID, "unsupported tuning values");
    d->tuning = *t;
    invalidate_graphs(d);
    return SLOSH_OK;
}

int slosh_get_tuning(const slosh_data *d, slosh_tuning *out) {
    if (!d || !out) return fail(SLOSH_E_INVALID, "null argument");
    *out = d->tuning;
    return SLOSH_OK;
}

// Times `steps` graph-replayed substeps of `c` under tuning `t` (best of two).
static int time_tuning(slosh_pipeline *pl, slosh_data *c, hipStream_t st, const slosh_tuning &t, uint32_t steps,
                       float *ms) {
    c->tuning = t;
    invalidate_graphs(c);
    TRY(slosh_launch_steps_graph(pl, c, st, 2));  // capture under this shape
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int rc = SLOSH_OK;
    for (int rep = 0; rep < 2 && rc == SLOSH_OK; rep++) {
        (void)hipEventRecord(e0, st);
        rc = slosh_launch_steps_graph(pl, c, st, steps);
        (void)hipEventRecord(e1, st);
        if (rc == SLOSH_OK && hipEventSynchronize(e1) == hipSuccess) {
            float v = 0.0f;
            if (hipEventElapsedTime(&v, e0, e1) == hipSuccess) best = std::min(best, v);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TRY(rc);
    TRY(slosh_check_errors(c, (slosh_stream)st));
    *ms = best;
    return SLOSH_OK;
}

int slosh_autotune(slosh_pipeline *pl, slosh_data *d, slosh_stream stream, uint32_t trial_steps, slosh_tuning *out) {
    if (!pl || !d) return fail(SLOSH_E_INVALID, "null pipeline or data");
    hipStream_t st = (hipStream_t)stream;
    if (d->n == 0) {
        if (out) *out = d->tuning;
        return SLOSH_OK;
    }
    const uint32_t steps = std::max<uint32_t>(2u, trial_steps + (trial_steps & 1u));
    // Trials run on a private copy so that `d` does not advance.
    RefArrays r;
    TRY(gather_reference(d, st, r));
    const slosh_particles p = r.view(d->batched != 0);
    slosh_data *c = nullptr;
    TRY(slosh_data_create(pl, &d->params, &p, d->pcap, d->bodies_host.data(), d->nbodies, d->mats_host.data(), d->h,
                          d->cap, &c));
    c->svd_f64 = d->svd_f64;
    if (d->mesh_vtx) {
        std::vector<float> mv(4 * d->mesh_nv);
        std::vector<uint32_t> mi(d->mesh_ni);
        if (d->mesh_nv) HIP_TRY(hipMemcpy(mv.data(), d->mesh_vtx, 16 * d->mesh_nv, hipMemcpyDeviceToHost));
        if (d->mesh_ni) HIP_TRY(hipMemcpy(mi.data(), d->mesh_idx, 4 * d->mesh_ni, hipMemcpyDeviceToHost));
        const int rm = slosh_data_set_collision_mesh(c, mv.data(), d->mesh_nv, mi.data(), d->mesh_ni);
        if (rm != SLOSH_OK) {
            slosh_data_destroy(c);
            return rm;
        }
    }
    slosh_tuning best = d->tuning;
    float best_ms = 0.0f;
    int rc = run_step(pl, c, st, nullptr, nullptr, nullptr);
    if (rc == SLOSH_OK) rc = time_tuning(pl, c, st, best, steps, &best_ms);
    // Coordinate search: P2G staging chunk, G2P workgroup size, persistent width.
    const uint32_t chunks[] = {192u, 256u, 384u, 512u};
    const uint32_t threads[] = {128u, 256u};
    const uint32_t widths[] = {2048u, 4096u, 8192u, 16384u, 32768u};
    auto consider = [&](slosh_tuning t) {
        if (rc != SLOSH_OK || std::memcmp(&t, &best, sizeof t) == 0) return;
        float ms = 0.0f;
        rc = time_tuning(pl, c, st, t, steps, &ms);
        if (rc == SLOSH_OK && ms < best_ms) {
            best = t;
            best_ms = ms;
        }
    };
    for (uint32_t v : chunks) {
        slosh_tuning t = best;
        t.p2g_chunk = v;
        consider(t);
    }
    for (uint32_t v : threads) {
        slosh_tuning t = best;
        t.g2p_threads = v;
        consider(t);
    }
    for (uint32_t v : widths) {
        if (v > 2 * d->cap) continue;  // wider than the block capacity adds nothing
        slosh_tuning t = best;
        t.block_wgs = v;
        consider(t);
    }
    std::string msg = g_last_error;
    slosh_data_destroy(c);
    if (rc != SLOSH_OK) {
        g_last_error = msg;
        return rc;
    }
    d->tuning = best;
    invalidate_graphs(d);
    if (out) *out = best;
    return SLOSH_OK;
}

int slosh_write_kinematics(slosh_data *d, slosh_stream stream, const slosh_kinematics *kin) {
    if (!d || (!kin && d->n)) return fail(SLOSH_E_INVALID, "null argument");
    hipStream_t st = (hipStream_t)stream;
    Snapshot s;
    TRY(download_current(d, st, s));
    const size_t n = d->n, pc = d->pcap;
    bool any_force = false;
    for (size_t i = 0; i < n; i++) {
        const slosh_kinematics &k = kin[s.ids[i]];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) s.C[(size_t)(r * 3 + c) * n + i] = k.affine.m[r][c];
        for (int c = 0; c < 3; c++) {
            s.vm[4 * i + c] = k.velocity[c];
            s.fdt[(size_t)c * n + i] = k.force_dt[c];
            if (k.force_dt[c] != 0.0f) any_force = true;
        }
        s.vm[4 * i + 3] = k.mass;
        s.vgd[i] = k.vel_grad_det;
        uint32_t w = f2u(s.pos[4 * i + 3]);
        s.pos[4 * i + 3] = u2f((w & ~1u) | (k.enabled ? 1u : 0u));
    }
    PBufs &p = d->pb[d->cur_p];
    if (n) {
        HIP_TRY(hipMemcpyAsync(p.pos, s.pos.data(), 16 * n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(p.vm, s.vm.data(), 16 * n, hipMemcpyHostToDevice, st));
        for (int k = 0; k < 9; k++) HIP_TRY(hipMemcpyAsync(p.C + k * pc, s.C.data() + k * n, 4 * n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(p.vgd, s.vgd.data(), 4 * n, hipMemcpyHostToDevice, st));
        for (int k = 0; k < 3; k++)
            HIP_TRY(hipMemcpyAsync(d->fdt + k * pc, s.fdt.data() + k * n, 4 * n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (any_force) d->has_force = 1;
    // The masses and enabled flags were all rewritten, and captured substeps may
    // carry the old uniform mass in their arguments.
    if (!d->masses_exposed) {
        d->mass_state = 0;
        note_masses(d, kin, n);
    }
    d->scene_masses_off = true;  // the per-scene masses are not re-derived
    invalidate_graphs(d);
    return SLOSH_OK;
}

int slosh_device_view_get(slosh_data *d, slosh_device_view *out) {
    if (!d || !out) return fail(SLOSH_E_INVALID, "null argument");
    if (!d->masses_exposed || d->mass_state != 2) {
        // The caller may write particle masses through the view: substeps captured
        // with the uniform mass in their arguments are retired.
        d->masses_exposed = true;
        d->mass_state = 2;
        invalidate_graphs(d);
    }
    PBufs &p = d->pb[d->cur_p];
    out->num_particles = d->n;
    out->particle_capacity = d->pcap;
    out->grid_capacity = d->cap;
    out->cell_width = d->h;
    out->positions = p.pos;
    out->vel_mass = p.vm;
    out->affine = p.C;
    out->def_grad = p.F;
    out->vel_grad_det = p.vgd;
    out->particle_ids = p.ids;
    out->force_dt = d->fdt;
    out->nodes = d->nodes;
    out->sorted_ids = d->perm;
    out->counters = ctr_set(d, d->cur_g);
    out->block_list = d->blist;
    return SLOSH_OK;
}

// ---- timestamps / streams / events -------------------------------------------
int slosh_timestamps_create(slosh_timestamps **out) {
    if (!out) return fail(SLOSH_E_INVALID, "null output");
    *out = new slosh_timestamps();
    return SLOSH_OK;
}

void slosh_timestamps_destroy(slosh_timestamps *ts) {
    if (!ts) return;
    for (auto &r : ts->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ts->pool) (void)hipEventDestroy(e);
    delete ts;
}

int slosh_timestamps_read(slosh_timestamps *ts, double *ms, uint32_t *count) {
    if (!ts) return fail(SLOSH_E_INVALID, "null timestamps");
    for (auto &r : ts->recs) {
        HIP_TRY(hipEventSynchronize(r.b));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
        ts->acc[r.pass] += t;
        ts->count[r.pass] += 1;
        ts->pool.push_back(r.a);
        ts->pool.push_back(r.b);
    }
    ts->recs.clear();
    for (int i = 0; i < SLOSH_NUM_PASSES; i++) {
        if (ms) ms[i] = ts->acc[i];
        if (count) count[i] = ts->count[i];
    }
    return SLOSH_OK;
}

void slosh_timestamps_reset(slosh_timestamps *ts) {
    if (!ts) return;
    for (int i = 0; i < SLOSH_NUM_PASSES; i++) {
        ts->acc[i] = 0.0;
        ts->count[i] = 0;
    }
}

int slosh_stream_create(slosh_stream *out) {
    if (!out) return fail(SLOSH_E_INVALID, "null output");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = (slosh_stream)s;
    return SLOSH_OK;
}

int slosh_stream_destroy(slosh_stream stream) {
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return SLOSH_OK;
}

int slosh_stream_synchronize(slosh_stream stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return SLOSH_OK;
}

int slosh_event_create(void **out) {
    if (!out) return fail(SLOSH_E_INVALID, "null output");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *out = (void *)e;
    return SLOSH_OK;
}

int slosh_event_destroy(void *event) {
    if (event) HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return SLOSH_OK;
}

int slosh_event_record(void *event, slosh_stream stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return SLOSH_OK;
}

int slosh_event_elapsed_ms(void *start, void *stop, float *ms) {
    HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SLOSH_OK;
}

int slosh_device_count(int *out) {
    if (!out) return fail(SLOSH_E_INVALID, "null output");
    HIP_TRY(hipGetDeviceCount(out));
    return SLOSH_OK;
}

}  // extern "C"
