/*
 * slosh_oracle.c -- CPU oracle for the slosh MPM substep. TEST INFRASTRUCTURE ONLY.
 *
 * A single-precision C restatement of the reference's substep, written from the
 * shader and host sources read as text. Each function cites the reference
 * file:line it restates. Atomic operations run sequentially in thread-index
 * order. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may
 * load this library; it is never part of the product path.
 *
 * Parity status of each part (how it is pinned) is recorded in the test suite:
 *  - keys / hash / block set / bucket counts / scan: closed forms + brute force
 *    (tests/test_oracle.py);
 *  - B-spline weights, P2G, G2P: partition of unity, momentum conservation,
 *    affine-field reproduction (C == A exactly);
 *  - stress models: closed forms (F=I, F=R, symmetry, hand cases for DP);
 *  - 3x3 SVD: reconstruction / orthogonality against numpy's LAPACK SVD.
 *  - collider projections: closed-form signed distances for ball, cuboid,
 *    capsule, cone and cylinder (tests/test_oracle_shapes.py); trimesh and
 *    heightfield projections by brute force over all triangles, a sloped-plane
 *    closed form and the convex inside test (tests/test_oracle_trimesh.py,
 *    tests/test_oracle_heightfield.py);
 *  - timestep bound: closed forms per material (tests/test_oracle.py);
 *  - Sim3 / quaternion helpers of the third-party `stensor` crate (absent from
 *    the reference tree): textbook formulas, parity unpinned beyond identities.
 */
#include "slosh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NONE 0xffffffffu
#define NUM_PRIMARY_SORT_BUCKETS 4u /* grid.slang:243 */
#define NUM_SORT_BUCKETS 10u        /* grid.slang:249 */
#define NUM_NBH_BLOCKS 7u           /* grid.slang:229 */
#define EXTRA_PARTICLE_MIN_SHIFT 2u /* grid.slang:256 */

typedef struct {
    uint32_t state;
    int32_t key[3];
    uint32_t value;
} hentry; /* GridHashMapEntry, grid.slang:105-115 */

typedef struct {
    int32_t vid[3];
    uint32_t first_particle;
    uint32_t num_particles_with_extras;
    uint32_t num_particles;
    uint32_t cursors[NUM_SORT_BUCKETS];
    uint32_t nbh[NUM_NBH_BLOCKS];
    uint32_t counts[NUM_SORT_BUCKETS]; /* copy of the bucket counts before they become cursors */
} header; /* ActiveBlockHeader, grid.slang:259-272 */

typedef struct {
    float normal[3];
    float point[3];
    uint32_t collider_id;
    uint32_t valid;
} ncoll; /* NodeCollision, collide.slang:51-65 */

struct so_state {
    size_t n;
    float h;
    uint32_t cap;
    so_sim_params params;
    float (*pos)[4];
    so_kinematics *kin;
    so_mat3 *F;
    so_props *props;
    so_model *models;
    size_t nb;
    so_body *bodies;
    so_bc *mats;
    int32_t imp[16][6]; /* IntegerImpulse linear/angular (rigid_impulses.slang:8-19) */
    hentry *hmap, *prev_hmap;
    uint32_t nblocks, prev_nblocks;
    header *hdr;
    uint32_t *scan;
    uint32_t *sorted;
    float (*nodes)[4];
    ncoll *coll, *prev_coll;
    int svd_mode;
    int overflow;
    so_mesh mesh; /* trimesh collision buffers (shapes_collision_vertices / _indices) */
    /* Per-particle sort scratch (filled in parallel, consumed in particle order). */
    int32_t (*pblk)[3];
    uint32_t (*passoc)[3];
    uint32_t *phid;
};

/* P2G formulation: 0 = block-local scatter (default), 1 = the shader's gather.
   Both give bit-identical node sums (see p2g_scatter). */
static int g_p2g_gather = 0;
void so_set_p2g_gather(int gather) { g_p2g_gather = gather; }

/* ------------------------------------------------------------------------- */
/* Small shader-semantics helpers (matrices are shader matrices, m[row][col]). */

typedef struct { float m[3][3]; } m3;

static m3 m3_zero(void) { m3 r; memset(&r, 0, sizeof r); return r; }
static m3 m3_ident(void) { m3 r = m3_zero(); r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0f; return r; }
static m3 m3_load(const so_mat3 *a) {
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a->m[i][j];
    return r;
}
static void m3_store(so_mat3 *a, m3 r) {
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) a->m[i][j] = r.m[i][j];
        a->m[i][3] = 0.0f;
    }
}
/* Slang mul(A, B) */
static m3 m3_mul(m3 a, m3 b) {
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
static m3 m3_t(m3 a) {
    m3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}
static m3 m3_scale(m3 a, float s) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.m[i][j] *= s;
    return a;
}
static m3 m3_add(m3 a, m3 b) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.m[i][j] += b.m[i][j];
    return a;
}
static m3 m3_sub(m3 a, m3 b) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.m[i][j] -= b.m[i][j];
    return a;
}
static float m3_det(m3 a) {
    return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
           a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
           a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
static float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float len3(const float *a) { return sqrtf(dot3(a, a)); }
static float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* ------------------------------------------------------------------------- */
/* Hash map + sparse grid, shaders/slosh/grid/grid.slang                       */

/* grid.slang:83-90 */
uint32_t so_pack_key(int32_t x, int32_t y, int32_t z) {
    return ((uint32_t)(x + 0x3ff) & 0x7ffu) | (((uint32_t)(y + 0x1ff) & 0x3ffu) << 11) |
           (((uint32_t)(z + 0x3ff) & 0x7ffu) << 21);
}

/* grid.slang:93-100 */
uint32_t so_hash(uint32_t k) {
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    return k;
}

/* grid.slang:118-174, sequential CAS. *full is set when no slot was found. */
static uint32_t insertion_index(hentry *hm, uint32_t cap, const int32_t key[3], int *full) {
    uint32_t pk = so_pack_key(key[0], key[1], key[2]);
    uint32_t slot = so_hash(pk) & (cap - 1u);
    for (uint32_t k = 0; k < cap; k++) {
        uint32_t old = hm[slot].state;
        if (old == NONE) {
            hm[slot].state = pk;
            memcpy(hm[slot].key, key, sizeof(int32_t) * 3);
            return slot;
        } else if (old == pk) {
            return NONE;
        }
        slot = (slot + 1u) % cap & (cap - 1u);
    }
    *full = 1;
    return NONE;
}

/* grid.slang:176-198 (bounded: the reference loops forever on a full table) */
static uint32_t find_block_header_id(const hentry *hm, uint32_t cap, const int32_t key[3]) {
    uint32_t pk = so_pack_key(key[0], key[1], key[2]);
    uint32_t slot = so_hash(pk) & (cap - 1u);
    for (uint32_t k = 0; k < cap; k++) {
        uint32_t st = hm[slot].state;
        if (st == pk) return hm[slot].value;
        if (st == NONE) break;
        slot = (slot + 1u) & (cap - 1u);
    }
    return NONE;
}

/* grid.slang:384-403 */
static void mark_block_as_active(so_state *s, const int32_t block[3]) {
    int full = 0;
    uint32_t slot = insertion_index(s->hmap, s->cap, block, &full);
    if (full) s->overflow = 1;
    if (slot != NONE) {
        uint32_t hid = s->nblocks++;
        header *hd = &s->hdr[hid];
        memcpy(hd->vid, block, sizeof(int32_t) * 3);
        hd->first_particle = 0;
        hd->num_particles_with_extras = 0;
        hd->num_particles = 0;
        for (uint32_t k = 0; k < NUM_SORT_BUCKETS; k++) hd->cursors[k] = 0;
        s->hmap[slot].value = hid;
    }
}

/* grid.slang:346-355 */
static void block_associated_to_point(float h, const float *pt, int32_t out[3]) {
    for (int i = 0; i < 3; i++) {
        float assoc_cell = rintf(pt[i] / h) - 1.0f;
        float assoc_block = floorf(assoc_cell / 4.0f);
        out[i] = (int32_t)assoc_block;
    }
}

/* particle.slang:79-87 */
static void associated_cell_index_in_block_off_by_one(float h, const float *pt, uint32_t out[3]) {
    for (int i = 0; i < 3; i++) {
        float assoc_cell = rintf(pt[i] / h) - 1.0f;
        float assoc_block = floorf(assoc_cell / 4.0f) * 4.0f;
        out[i] = (uint32_t)(assoc_cell - assoc_block);
    }
}

/* grid.slang:362-382: the 8 associated blocks, primary first. */
static const int32_t ASSOC_OFFSETS[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                                            {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

/* sort.slang:15-21 */
static uint32_t primary_sort_bucket(const uint32_t assoc[3]) { return assoc[2]; }

/* sort.slang:27-34 */
static uint32_t extra_sort_bucket(const uint32_t assoc[3], const int32_t bshift[3]) {
    int32_t local = (int32_t)assoc[2] - bshift[2] * 4;
    int32_t m = local > -2 ? local : -2;
    return NUM_PRIMARY_SORT_BUCKETS + (uint32_t)(m + 2);
}

/* sort.slang:38-42 */
static int extra_spills(const uint32_t assoc[3], const int32_t bshift[3]) {
    for (int i = 0; i < 3; i++) {
        uint32_t mask = assoc[i] >= EXTRA_PARTICLE_MIN_SHIFT ? 1u : 0u;
        uint32_t bs = (uint32_t)bshift[i];
        if (bs * mask != bs) return 0;
    }
    return 1;
}

/* prefix_sum.rs:97-109 semantics (exclusive scan, "as if a 0 was prepended"). */
void so_prefix_sum(uint32_t *v, size_t len) {
    uint32_t run = 0;
    for (size_t i = 0; i < len; i++) {
        uint32_t x = v[i];
        v[i] = run;
        run += x;
    }
}

/* GpuGrid::swap_buffers (grid.rs:513-517) + WgGrid::launch_sort (grid.rs:78-238). */
static void grid_sort(so_state *s) {
    const float h = s->h;
    /* swap_buffers: meta, hmap entries and node collisions ping-pong. */
    hentry *th = s->hmap; s->hmap = s->prev_hmap; s->prev_hmap = th;
    ncoll *tc = s->coll; s->coll = s->prev_coll; s->prev_coll = tc;
    uint32_t tn = s->nblocks; s->nblocks = s->prev_nblocks; s->prev_nblocks = tn;

    /* reset_hmap, grid.slang:200-218 */
    for (uint32_t i = 0; i < s->cap; i++) s->hmap[i].state = NONE;
    s->nblocks = 0;

    /* The associated block and cell of every particle (pure functions of the
       position, computed in parallel; the passes below read them in particle order). */
#pragma omp parallel for schedule(static)
    for (long id = 0; id < (long)s->n; id++) {
        block_associated_to_point(h, s->pos[id], s->pblk[id]);
        associated_cell_index_in_block_off_by_one(h, s->pos[id], s->passoc[id]);
    }
    /* touch_primary_blocks, sort.slang:50-67 (all particles: the intended
       semantics of the capped launch, see grid.rs:131-136), sequential in particle
       order. Inserting a key already present is a no-op, so a particle in the same
       block as the previous one is skipped. */
    for (size_t id = 0; id < s->n; id++) {
        const int32_t *b = s->pblk[id];
        if (id > 0 && memcmp(b, s->pblk[id - 1], sizeof(int32_t) * 3) == 0) continue;
        mark_block_as_active(s, b);
    }
    /* capture_num_active_blocks, grid.slang:428-436 */
    uint32_t num_base_blocks = s->nblocks;
    /* touch_neighbor_blocks, sort.slang:75-92 */
    for (uint32_t id = 0; id < num_base_blocks; id++) {
        int32_t vid[3] = {s->hdr[id].vid[0], s->hdr[id].vid[1], s->hdr[id].vid[2]};
        for (int i = 1; i < 8; i++) {
            int32_t b[3] = {vid[0] + ASSOC_OFFSETS[i][0], vid[1] + ASSOC_OFFSETS[i][1], vid[2] + ASSOC_OFFSETS[i][2]};
            mark_block_as_active(s, b);
        }
    }
    /* update_nbh_block_ids, sort.slang:100-117 */
#pragma omp parallel for schedule(static)
    for (long id = 0; id < (long)s->nblocks; id++) {
        for (uint32_t k = 0; k < NUM_NBH_BLOCKS; k++) {
            int32_t b[3] = {s->hdr[id].vid[0] + ASSOC_OFFSETS[k + 1][0], s->hdr[id].vid[1] + ASSOC_OFFSETS[k + 1][1],
                            s->hdr[id].vid[2] + ASSOC_OFFSETS[k + 1][2]};
            s->hdr[id].nbh[k] = find_block_header_id(s->hmap, s->cap, b);
        }
    }
    /* Header id of every particle's block (read-only lookups, in parallel). */
#pragma omp parallel for schedule(static)
    for (long id = 0; id < (long)s->n; id++) s->phid[id] = find_block_header_id(s->hmap, s->cap, s->pblk[id]);
    /* update_block_particle_count, sort.slang:208-246 */
    for (size_t id = 0; id < s->n; id++) {
        const uint32_t *assoc = s->passoc[id];
        uint32_t hid0 = s->phid[id];
        if (hid0 == NONE) { s->overflow = 1; continue; }
        s->hdr[hid0].cursors[primary_sort_bucket(assoc)] += 1u;
        for (int i = 1; i < 8; i++) {
            const int32_t *bshift = ASSOC_OFFSETS[i];
            if (extra_spills(assoc, bshift)) {
                uint32_t bi = s->hdr[hid0].nbh[i - 1];
                if (bi == NONE) { s->overflow = 1; continue; }
                s->hdr[bi].cursors[extra_sort_bucket(assoc, bshift)] += 1u;
            }
        }
    }
    /* copy_particles_len_to_scan_value, sort.slang:248-267 */
    for (uint32_t id = 0; id < s->nblocks; id++) {
        uint32_t total = 0;
        for (uint32_t k = 0; k < NUM_SORT_BUCKETS; k++) {
            total += s->hdr[id].cursors[k];
            s->hdr[id].counts[k] = s->hdr[id].cursors[k];
        }
        s->scan[id] = total;
    }
    /* prefix_sum over the whole buffer (prefix_sum.slang:7-101). Entries past
       num_active_blocks only influence outputs past it, which are never read. */
    so_prefix_sum(s->scan, s->nblocks);
    /* copy_scan_values_to_first_particles_and_prepare_for_finalize, sort.slang:269-302 */
    for (uint32_t id = 0; id < s->nblocks; id++) {
        uint32_t first = s->scan[id];
        header *hd = &s->hdr[id];
        hd->first_particle = first;
        uint32_t running = first;
        for (uint32_t k = 0; k < NUM_SORT_BUCKETS; k++) {
            if (k == NUM_PRIMARY_SORT_BUCKETS) hd->num_particles = running - first;
            uint32_t count = hd->cursors[k];
            hd->cursors[k] = running;
            running += count;
        }
        hd->num_particles_with_extras = running - first;
    }
    /* reset, grid.slang:455-485 (momentum lane; CPIC lanes are never read). */
    memset(s->nodes, 0, sizeof(float) * 4 * (size_t)s->nblocks * 64u);
    /* finalize_particles_sort, sort.slang:304-357 (the per-node linked lists it also
       builds are read only by kernels launch_step never dispatches). */
    for (size_t id = 0; id < s->n; id++) {
        const uint32_t *assoc = s->passoc[id];
        uint32_t hid0 = s->phid[id];
        if (hid0 == NONE) continue;
        uint32_t slot0 = s->hdr[hid0].cursors[primary_sort_bucket(assoc)]++;
        s->sorted[slot0] = (uint32_t)id;
        for (int i = 1; i < 8; i++) {
            const int32_t *bshift = ASSOC_OFFSETS[i];
            if (extra_spills(assoc, bshift)) {
                uint32_t bi = s->hdr[hid0].nbh[i - 1];
                if (bi == NONE) continue;
                uint32_t sl = s->hdr[bi].cursors[extra_sort_bucket(assoc, bshift)]++;
                s->sorted[sl] = (uint32_t)id;
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Quadratic B-spline, shaders/slosh/grid/kernel.slang                          */

/* kernel.slang:168-174 */
static float k_eval(float x) {
    float x_abs = fabsf(x);
    float part1 = 0.75f - x_abs * x_abs;
    float part2 = 0.5f * (1.5f - x_abs) * (1.5f - x_abs);
    return x_abs < 0.5f ? part1 : (x_abs < 1.5f ? part2 : 0.0f);
}

/* kernel.slang:160-166 */
static void k_eval_all(float x, float out[3]) {
    out[0] = 0.5f * (1.5f - x) * (1.5f - x);
    out[1] = 0.75f - (x - 1.0f) * (x - 1.0f);
    out[2] = 0.5f * (x - 0.5f) * (x - 0.5f);
}

/* kernel.slang:120-152 */
static const uint32_t NBH_SHIFTS[27][3] = {
    {2, 2, 2}, {2, 0, 2}, {2, 1, 2}, {0, 2, 2}, {0, 0, 2}, {0, 1, 2}, {1, 2, 2}, {1, 0, 2}, {1, 1, 2},
    {2, 2, 0}, {2, 0, 0}, {2, 1, 0}, {0, 2, 0}, {0, 0, 0}, {0, 1, 0}, {1, 2, 0}, {1, 0, 0}, {1, 1, 0},
    {2, 2, 1}, {2, 0, 1}, {2, 1, 1}, {0, 2, 1}, {0, 0, 1}, {0, 1, 1}, {1, 2, 1}, {1, 0, 1}, {1, 1, 1}};

/* ------------------------------------------------------------------------- */
/* P2G, shaders/slosh/solver/p2g_scatter_style.slang:41-155                     */

static void p2g_gather(so_state *s) {
    const float cell_width = s->h;
    const float inv_cell_width = 1.0f / cell_width;
#pragma omp parallel for schedule(dynamic, 4)
    for (long bid = 0; bid < (long)s->nblocks; bid++) {
        const header *hd = &s->hdr[bid];
        uint32_t first_particle = hd->first_particle;
        uint32_t last_particle = first_particle + hd->num_particles_with_extras;
        uint32_t primaries_end = first_particle + hd->num_particles;
        float sh_pos[64][3], sh_mom[64][3], sh_mass[64];
        m3 sh_aff[64];
        int sh_zkey[64];
        float acc[64][4];
        memset(acc, 0, sizeof acc);
        for (uint32_t chunk_base = first_particle; chunk_base < last_particle; chunk_base += 64u) {
            uint32_t chunk_len = last_particle - chunk_base < 64u ? last_particle - chunk_base : 64u;
            for (uint32_t t = 0; t < chunk_len; t++) {
                uint32_t pid = s->sorted[chunk_base + t];
                const so_kinematics *kin = &s->kin[pid];
                const float *pos = s->pos[pid];
                for (int i = 0; i < 3; i++) {
                    sh_pos[t][i] = pos[i];
                    sh_mom[t][i] = kin->velocity[i] * kin->mass + kin->force_dt[i];
                }
                sh_mass[t] = kin->mass;
                sh_aff[t] = m3_load(&kin->affine);
                float assoc_z = rintf(pos[2] / cell_width) - 1.0f;
                int zk = (int)assoc_z - hd->vid[2] * 4;
                sh_zkey[t] = zk > -2 ? zk : -2;
            }
            int zmin = sh_zkey[0];
            int zmax = sh_zkey[chunk_len - 1u];
            if (primaries_end > chunk_base && primaries_end < chunk_base + chunk_len) {
                uint32_t b = primaries_end - chunk_base;
                if (sh_zkey[b] < zmin) zmin = sh_zkey[b];
                if (sh_zkey[b - 1u] > zmax) zmax = sh_zkey[b - 1u];
            }
            for (uint32_t tid = 0; tid < 64u; tid++) {
                int lc[3] = {(int)(tid % 4u), (int)((tid / 4u) % 4u), (int)(tid / 16u)};
                float cell_pos[3];
                for (int i = 0; i < 3; i++) cell_pos[i] = (float)(hd->vid[i] * 4 + lc[i]) * cell_width;
                int node_slab = lc[2];
                uint32_t culled_len = (node_slab >= zmin && node_slab <= zmax + 2) ? chunk_len : 0u;
                for (uint32_t p = 0; p < culled_len; p++) {
                    float dpt[3] = {cell_pos[0] - sh_pos[p][0], cell_pos[1] - sh_pos[p][1], cell_pos[2] - sh_pos[p][2]};
                    float weight = k_eval(dpt[0] * inv_cell_width) * k_eval(dpt[1] * inv_cell_width) *
                                   k_eval(dpt[2] * inv_cell_width);
                    if (weight != 0.0f) {
                        float mom[3];
                        for (int j = 0; j < 3; j++)
                            mom[j] = dpt[0] * sh_aff[p].m[0][j] + dpt[1] * sh_aff[p].m[1][j] +
                                     dpt[2] * sh_aff[p].m[2][j] + sh_mom[p][j];
                        acc[tid][0] += mom[0] * weight;
                        acc[tid][1] += mom[1] * weight;
                        acc[tid][2] += mom[2] * weight;
                        acc[tid][3] += sh_mass[p] * weight;
                    }
                }
            }
        }
        memcpy(s->nodes[(size_t)bid * 64u], acc, sizeof acc);
    }
}

/* The same P2G as a block-local scatter. In the gather above, node `tid` of a
   block visits the chunk's particles in list order and adds the contribution of
   each one whose weight k(dx) k(dy) k(dz) is non-zero, unless the z-slab culling
   of the chunk excludes the node. Here each particle, in list order, adds its
   contribution to every node of the block that the culling keeps and whose weight
   is non-zero. The per-axis factors depend only on that axis's node coordinate,
   so they are evaluated once per particle for the block's 4 coordinates of each
   axis; every node then receives the same terms, computed by the same operations
   (no contraction), in the same order: the node sums are bit-identical, at ~27
   instead of 64 kernel evaluations per listed particle. */
static void p2g_scatter(so_state *s) {
    const float cell_width = s->h;
    const float inv_cell_width = 1.0f / cell_width;
#pragma omp parallel for schedule(dynamic, 4)
    for (long bid = 0; bid < (long)s->nblocks; bid++) {
        const header *hd = &s->hdr[bid];
        uint32_t first_particle = hd->first_particle;
        uint32_t last_particle = first_particle + hd->num_particles_with_extras;
        uint32_t primaries_end = first_particle + hd->num_particles;
        float node_coord[3][4];
        for (int i = 0; i < 3; i++)
            for (int l = 0; l < 4; l++) node_coord[i][l] = (float)(hd->vid[i] * 4 + l) * cell_width;
        float acc[64][4];
        memset(acc, 0, sizeof acc);
        int sh_zkey[64];
        for (uint32_t chunk_base = first_particle; chunk_base < last_particle; chunk_base += 64u) {
            uint32_t chunk_len = last_particle - chunk_base < 64u ? last_particle - chunk_base : 64u;
            for (uint32_t t = 0; t < chunk_len; t++) {
                const float *pos = s->pos[s->sorted[chunk_base + t]];
                float assoc_z = rintf(pos[2] / cell_width) - 1.0f;
                int zk = (int)assoc_z - hd->vid[2] * 4;
                sh_zkey[t] = zk > -2 ? zk : -2;
            }
            int zmin = sh_zkey[0];
            int zmax = sh_zkey[chunk_len - 1u];
            if (primaries_end > chunk_base && primaries_end < chunk_base + chunk_len) {
                uint32_t b = primaries_end - chunk_base;
                if (sh_zkey[b] < zmin) zmin = sh_zkey[b];
                if (sh_zkey[b - 1u] > zmax) zmax = sh_zkey[b - 1u];
            }
            for (uint32_t t = 0; t < chunk_len; t++) {
                uint32_t pid = s->sorted[chunk_base + t];
                const so_kinematics *kin = &s->kin[pid];
                const float *pos = s->pos[pid];
                float mom0[3];
                for (int i = 0; i < 3; i++) mom0[i] = kin->velocity[i] * kin->mass + kin->force_dt[i];
                const float mass = kin->mass;
                const m3 aff = m3_load(&kin->affine);
                float d[3][4], w[3][4];
                for (int i = 0; i < 3; i++)
                    for (int l = 0; l < 4; l++) {
                        d[i][l] = node_coord[i][l] - pos[i];
                        w[i][l] = k_eval(d[i][l] * inv_cell_width);
                    }
                for (int lz = 0; lz < 4; lz++) {
                    if (!(lz >= zmin && lz <= zmax + 2) || w[2][lz] == 0.0f) continue;
                    for (int ly = 0; ly < 4; ly++) {
                        if (w[1][ly] == 0.0f) continue;
                        for (int lx = 0; lx < 4; lx++) {
                            if (w[0][lx] == 0.0f) continue;
                            float weight = w[0][lx] * w[1][ly] * w[2][lz];
                            if (weight == 0.0f) continue;
                            const float dpt[3] = {d[0][lx], d[1][ly], d[2][lz]};
                            float mom[3];
                            for (int j = 0; j < 3; j++)
                                mom[j] = dpt[0] * aff.m[0][j] + dpt[1] * aff.m[1][j] + dpt[2] * aff.m[2][j] + mom0[j];
                            float *a = acc[lx + 4 * ly + 16 * lz];
                            a[0] += mom[0] * weight;
                            a[1] += mom[1] * weight;
                            a[2] += mom[2] * weight;
                            a[3] += mass * weight;
                        }
                    }
                }
            }
        }
        memcpy(s->nodes[(size_t)bid * 64u], acc, sizeof acc);
    }
}

static void p2g(so_state *s) {
    if (g_p2g_gather) p2g_gather(s);
    else p2g_scatter(s);
}

/* ------------------------------------------------------------------------- */
/* Rigid shapes / poses (shaders/slosh/rbd) and collisions                      */

static void quat_rotate(const float q[4], const float v[3], float out[3]) {
    /* v' = v + w t + u x t, t = 2 u x v (unit quaternion q = (u, w)). */
    float t[3] = {2.0f * (q[1] * v[2] - q[2] * v[1]), 2.0f * (q[2] * v[0] - q[0] * v[2]),
                  2.0f * (q[0] * v[1] - q[1] * v[0])};
    out[0] = v[0] + q[3] * t[0] + (q[1] * t[2] - q[2] * t[1]);
    out[1] = v[1] + q[3] * t[1] + (q[2] * t[0] - q[0] * t[2]);
    out[2] = v[2] + q[3] * t[2] + (q[0] * t[1] - q[1] * t[0]);
}
static void quat_inv_rotate(const float q[4], const float v[3], float out[3]) {
    float qc[4] = {-q[0], -q[1], -q[2], q[3]};
    quat_rotate(qc, v, out);
}
/* Sim3::mul_pt (stensor, not vendored): rotation * (p * scale) + translation. */
static void sim_mul_pt(const so_sim3 *p, const float pt[3], float out[3]) {
    float sp[3] = {pt[0] * p->translation_scale[3], pt[1] * p->translation_scale[3], pt[2] * p->translation_scale[3]};
    quat_rotate(p->rotation, sp, out);
    for (int i = 0; i < 3; i++) out[i] += p->translation_scale[i];
}
/* Sim3::inv_mul_pt: rotation^-1 * (p - translation) / scale. */
static void sim_inv_mul_pt(const so_sim3 *p, const float pt[3], float out[3]) {
    float d[3] = {pt[0] - p->translation_scale[0], pt[1] - p->translation_scale[1], pt[2] - p->translation_scale[2]};
    quat_inv_rotate(p->rotation, d, out);
    for (int i = 0; i < 3; i++) out[i] /= p->translation_scale[3];
}

/* cuboid.slang:14-22 */
static void cuboid_project_local_point(const float he[3], const float pt[3], float out[3]) {
    for (int i = 0; i < 3; i++) {
        float mins_pt = -he[i] - pt[i];
        float pt_maxs = pt[i] - he[i];
        float shift = fmaxf(mins_pt, 0.0f) - fmaxf(pt_maxs, 0.0f);
        out[i] = pt[i] + shift;
    }
}

/* cuboid.slang:34-63 */
static int cuboid_project_local_point_on_boundary(const float he[3], const float pt[3], float out[3]) {
    float out_proj[3];
    cuboid_project_local_point(he, pt, out_proj);
    float pt_sgn[3], diff[3];
    for (int i = 0; i < 3; i++) {
        float swz = fsign(pt[i]);
        pt_sgn[i] = swz + (fabsf(swz) - 1.0f);
        diff[i] = he[i] - pt_sgn[i] * pt[i];
    }
    int pick_x = diff[0] <= diff[1] && diff[0] <= diff[2];
    int pick_y = diff[1] <= diff[0] && diff[1] <= diff[2];
    float pen[3] = {0.0f, 0.0f, 0.0f};
    if (pick_x) pen[0] = diff[0] * pt_sgn[0];
    else if (pick_y) pen[1] = diff[1] * pt_sgn[1];
    else pen[2] = diff[2] * pt_sgn[2];
    int is_inside = pt[0] == out_proj[0] && pt[1] == out_proj[1] && pt[2] == out_proj[2];
    for (int i = 0; i < 3; i++) out[i] = is_inside ? pt[i] + pen[i] : out_proj[i];
    return is_inside;
}

/* ball.slang:39-50 */
static int ball_project_local_point_on_boundary(float radius, const float pt[3], float out[3]) {
    float dist = len3(pt);
    if (dist != 0.0f) {
        float k = radius / dist;
        for (int i = 0; i < 3; i++) out[i] = pt[i] * k;
    } else {
        out[0] = 0.0f; out[1] = radius; out[2] = 0.0f;
    }
    return dist <= radius;
}

/* segment.slang:10-28: closest point of segment [a, b] to pt. */
static void segment_project_local_point(const float a[3], const float b[3], const float pt[3], float out[3]) {
    float ab[3], ap[3];
    for (int i = 0; i < 3; i++) {
        ab[i] = b[i] - a[i];
        ap[i] = pt[i] - a[i];
    }
    const float ab_ap = dot3(ab, ap), sqnab = dot3(ab, ab);
    if (ab_ap <= 0.0f) {
        memcpy(out, a, sizeof(float) * 3);
    } else if (ab_ap >= sqnab) {
        memcpy(out, b, sizeof(float) * 3);
    } else {
        const float u = ab_ap / sqnab;
        for (int i = 0; i < 3; i++) out[i] = a[i] + ab[i] * u;
    }
}

/* capsule.slang:77-100: the first vector of the orthonormal basis built from unit v. */
static void any_orthogonal_vector(const float v[3], float out[3]) {
    const float sign = v[2] >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + v[2]);
    const float b = v[0] * v[1] * a;
    out[0] = 1.0f + sign * v[0] * v[0] * a;
    out[1] = sign * b;
    out[2] = -sign * v[0];
}

/* capsule.slang:41-58 (segment a..b, radius) */
static int capsule_project_local_point_on_boundary(const float a[3], const float b[3], float radius, const float pt[3],
                                                   float out[3]) {
    float on_axis[3], d[3];
    segment_project_local_point(a, b, pt, on_axis);
    for (int i = 0; i < 3; i++) d[i] = pt[i] - on_axis[i];
    const float dist = len3(d);
    if (dist > 0.0f) {
        const float k = radius / dist;
        for (int i = 0; i < 3; i++) out[i] = on_axis[i] + d[i] * k;
        return dist <= radius;
    }
    /* On the axis: any direction orthogonal to it. */
    float ax[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float len = len3(ax);
    for (int i = 0; i < 3; i++) ax[i] /= len;
    float dir[3];
    any_orthogonal_vector(ax, dir);
    for (int i = 0; i < 3; i++) out[i] = on_axis[i] + dir[i] * radius;
    return 1;
}

/* Unit direction of (x, z) from the y axis, (1, 0) on the axis; returns |(x, z)|. */
static float planar_dir(const float pt[3], float *dx, float *dz) {
    const float pd = sqrtf(pt[0] * pt[0] + pt[2] * pt[2]);
    if (pd > 0.0f) {
        *dx = pt[0] / pd;
        *dz = pt[2] / pd;
    } else {
        *dx = 1.0f;
        *dz = 0.0f;
    }
    return pd;
}

/* cylinder.slang:68-128 (axis y, half height hh, radius r) */
static int cylinder_project_local_point_on_boundary(float hh, float r, const float pt[3], float out[3]) {
    float dx, dz;
    const float pd = planar_dir(pt, &dx, &dz);
    const float px = dx * r, pz = dz * r;
    if (pt[1] >= -hh && pt[1] <= hh && pd <= r) {
        const float to_top = hh - pt[1], to_bottom = pt[1] - (-hh), to_side = r - pd;
        if (to_top < to_bottom && to_top < to_side) {
            out[0] = pt[0]; out[1] = hh; out[2] = pt[2];
        } else if (to_bottom < to_top && to_bottom < to_side) {
            out[0] = pt[0]; out[1] = -hh; out[2] = pt[2];
        } else {
            out[0] = px; out[1] = pt[1]; out[2] = pz;
        }
        return 1;
    }
    if (pt[1] > hh || pt[1] < -hh) {
        const float y = pt[1] > hh ? hh : -hh;
        if (pd <= r) {
            out[0] = pt[0]; out[1] = y; out[2] = pt[2];
        } else {
            out[0] = px; out[1] = y; out[2] = pz;
        }
    } else {
        out[0] = px; out[1] = pt[1]; out[2] = pz;
    }
    return 0;
}

static void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* cone.slang:61-120 (axis y, apex at +hh, base disc of radius r at -hh) */
static int cone_project_local_point_on_boundary(float hh, float r, const float pt[3], float out[3]) {
    float dx, dz;
    const float pd = planar_dir(pt, &dx, &dz);
    const float on_base[3] = {pt[0], -hh, pt[2]};
    if (pt[1] < -hh && pd <= r) {
        memcpy(out, on_base, sizeof on_base);
        return 0;
    }
    const float rim[3] = {dx * r, -hh, dz * r};
    const float apex[3] = {0.0f, hh, 0.0f};
    const float side[3] = {rim[0] - apex[0], rim[1] - apex[1], rim[2] - apex[2]};
    float proj[3];
    segment_project_local_point(apex, rim, pt, proj);
    const float apex_to_base[3] = {0.0f, -2.0f * hh, 0.0f};
    const float ap[3] = {pt[0] - apex[0], pt[1] - apex[1], pt[2] - apex[2]};
    float c1[3], c2[3];
    cross3(side, ap, c1);
    cross3(side, apex_to_base, c2);
    if (pt[1] >= -hh && pt[1] <= hh && dot3(c1, c2) >= 0.0f) {
        const float to_proj[3] = {proj[0] - pt[0], proj[1] - pt[1], proj[2] - pt[2]};
        const float to_base[3] = {on_base[0] - pt[0], on_base[1] - pt[1], on_base[2] - pt[2]};
        if (dot3(to_proj, to_proj) > dot3(to_base, to_base)) memcpy(out, on_base, sizeof on_base);
        else memcpy(out, proj, sizeof proj);
        return 1;
    }
    memcpy(out, proj, sizeof proj);
    return 0;
}

/* ---- triangle meshes (collision/trimesh.slang) ---- */

enum { FEAT_VERTEX = 0, FEAT_EDGE = 1, FEAT_FACE = 2, FEAT_SOLID = 3 };

/* triangle_project_local_point_and_get_location (trimesh.slang:304-377): closest
 * point of triangle (a, b, c) to pt and its feature (vertex 0..2 = a, b, c; edge
 * 0 = ab, 1 = bc, 2 = ca; face). */
static void triangle_project(const float a[3], const float b[3], const float c[3], const float pt[3], float out[3],
                             uint32_t *feat, uint32_t *id) {
    float ab[3], ac[3], ap[3], bp[3], cp[3], bc[3];
    for (int i = 0; i < 3; i++) {
        ab[i] = b[i] - a[i];
        ac[i] = c[i] - a[i];
        ap[i] = pt[i] - a[i];
    }
    const float ab_ap = dot3(ab, ap), ac_ap = dot3(ac, ap);
    if (ab_ap <= 0.0f && ac_ap <= 0.0f) {
        memcpy(out, a, 12); *feat = FEAT_VERTEX; *id = 0;
        return;
    }
    for (int i = 0; i < 3; i++) bp[i] = pt[i] - b[i];
    const float ab_bp = dot3(ab, bp), ac_bp = dot3(ac, bp);
    if (ab_bp >= 0.0f && ac_bp <= ab_bp) {
        memcpy(out, b, 12); *feat = FEAT_VERTEX; *id = 1;
        return;
    }
    for (int i = 0; i < 3; i++) cp[i] = pt[i] - c[i];
    const float ab_cp = dot3(ab, cp), ac_cp = dot3(ac, cp);
    if (ac_cp >= 0.0f && ab_cp <= ac_cp) {
        memcpy(out, c, 12); *feat = FEAT_VERTEX; *id = 2;
        return;
    }
    for (int i = 0; i < 3; i++) bc[i] = c[i] - b[i];
    /* stable_check_edges_voronoi (trimesh.slang:257-302) */
    float n[3], t[3];
    cross3(ab, ac, n);
    cross3(ab, ap, t);
    const float vc = dot3(n, t);
    if (vc < 0.0f && ab_ap >= 0.0f && ab_bp <= 0.0f) {
        const float v = ab_ap / dot3(ab, ab);
        for (int i = 0; i < 3; i++) out[i] = a[i] + ab[i] * v;
        *feat = FEAT_EDGE; *id = 0;
        return;
    }
    cross3(ac, cp, t);
    const float vb = -dot3(n, t);
    if (vb < 0.0f && ac_ap >= 0.0f && ac_cp <= 0.0f) {
        const float w = ac_ap / dot3(ac, ac);
        for (int i = 0; i < 3; i++) out[i] = a[i] + ac[i] * w;
        *feat = FEAT_EDGE; *id = 2;
        return;
    }
    cross3(bc, bp, t);
    const float va = dot3(n, t);
    if (va < 0.0f && ac_bp - ab_bp >= 0.0f && ab_cp - ac_cp >= 0.0f) {
        const float w = dot3(bc, bp) / dot3(bc, bc);
        for (int i = 0; i < 3; i++) out[i] = b[i] + bc[i] * w;
        *feat = FEAT_EDGE; *id = 1;
        return;
    }
    if (va + vb + vc != 0.0f) {
        const float denom = 1.0f / (va + vb + vc);
        const float v = vb * denom, w = vc * denom;
        for (int i = 0; i < 3; i++) out[i] = a[i] + ab[i] * v + ac[i] * w;
        *feat = FEAT_FACE; *id = 0;
        return;
    }
    memcpy(out, pt, 12); /* degenerate triangle (trimesh.slang:372-376) */
    *feat = FEAT_SOLID; *id = 0;
}

static const float *mesh_v(const so_mesh *m, uint32_t i) { return m->vtx + 4 * (size_t)i; }

/* TriMesh::project_local_point (trimesh.slang:111-154): the closest triangle within
 * max_dist by the stackless BVH walk, inside by the pseudo-normal of the feature
 * the point projects on. Returns found. */
static int trimesh_project_local_point(const so_mesh *m, const so_shape *sh, const float pt[3], float max_dist,
                                       float out[3], int *inside) {
    uint32_t vroot, iroot, nodes, ntri, nvert;
    memcpy(&vroot, &sh->a[0], 4);
    memcpy(&iroot, &sh->a[1], 4);
    memcpy(&nodes, &sh->a[2], 4);
    memcpy(&ntri, &sh->b[0], 4);
    memcpy(&nvert, &sh->b[1], 4);
    (void)ntri;
    float best = max_dist, best_p[3] = {pt[0], pt[1], pt[2]};
    uint32_t best_tri = 0, best_feat = FEAT_SOLID, best_id = 0;
    int found = 0;
    const uint32_t *tidx = m->idx + iroot + 3 * (size_t)nodes;
    const uint32_t vbase = vroot + 2 * nodes;
    uint32_t cur = 0;
    while (cur < nodes) {
        const uint32_t *node = m->idx + iroot + 3 * (size_t)cur;
        if (node[0] == 0xffffffffu) {
            const uint32_t tri = node[2];
            float p[3];
            uint32_t feat, id;
            triangle_project(mesh_v(m, vbase + tidx[3 * tri]), mesh_v(m, vbase + tidx[3 * tri + 1]),
                             mesh_v(m, vbase + tidx[3 * tri + 2]), pt, p, &feat, &id);
            const float d[3] = {p[0] - pt[0], p[1] - pt[1], p[2] - pt[2]};
            const float dist = len3(d);
            if (dist <= best) {
                best = dist;
                memcpy(best_p, p, 12);
                best_tri = tri; best_feat = feat; best_id = id;
                found = 1;
            }
            cur = node[1];
        } else {
            const float *mn = mesh_v(m, vroot + 2 * cur), *mx = mesh_v(m, vroot + 2 * cur + 1);
            float d[3];
            for (int i = 0; i < 3; i++) d[i] = (pt[i] + (fmaxf(mn[i] - pt[i], 0.0f) - fmaxf(pt[i] - mx[i], 0.0f))) - pt[i];
            cur = len3(d) < best ? node[0] : node[1];
        }
    }
    /* TriMesh::pseudo_normal (trimesh.slang:83-109) */
    float pn[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t *tv = tidx + 3 * (size_t)best_tri;
    if (best_feat == FEAT_VERTEX) {
        memcpy(pn, mesh_v(m, vbase + tv[best_id] + nvert), 12);
    } else if (best_feat == FEAT_EDGE) {
        memcpy(pn, mesh_v(m, vbase + 2 * nvert + 3 * best_tri + best_id), 12);
    } else {
        const float *a = mesh_v(m, vbase + tv[0]), *b = mesh_v(m, vbase + tv[1]), *c = mesh_v(m, vbase + tv[2]);
        const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        cross3(ab, ac, pn);
    }
    const float dp[3] = {pt[0] - best_p[0], pt[1] - best_p[1], pt[2] - best_p[2]};
    *inside = dot3(pn, dp) <= 0.0f;
    memcpy(out, best_p, 12);
    return found;
}

/* shape.slang:180-201 and the trimesh branch of collide.slang:120-130: returns
 * handled (0 for polylines, and for trimeshes with no triangle within max_dist). */
static int shape_project_point_on_boundary(const so_shape *sh, const so_sim3 *pose, const float pt[3],
                                           const so_mesh *mesh, float max_dist, float out[3], int *inside) {
    uint32_t ty;
    memcpy(&ty, &sh->a[3], 4);
    float lp[3], lproj[3];
    sim_inv_mul_pt(pose, pt, lp);
    if (ty == 1u) {
        *inside = cuboid_project_local_point_on_boundary(sh->a, lp, lproj);
    } else if (ty == 0u) {
        *inside = ball_project_local_point_on_boundary(sh->a[0], lp, lproj);
    } else if (ty == 2u) {
        *inside = capsule_project_local_point_on_boundary(sh->a, sh->b, sh->b[3], lp, lproj);
    } else if (ty == 3u) {
        *inside = cone_project_local_point_on_boundary(sh->a[0], sh->a[1], lp, lproj);
    } else if (ty == 4u) {
        *inside = cylinder_project_local_point_on_boundary(sh->a[0], sh->a[1], lp, lproj);
    } else if (ty == 6u) {
        if (!mesh || !mesh->vtx || !trimesh_project_local_point(mesh, sh, lp, max_dist, lproj, inside)) return 0;
    } else if (ty == 5u) {
        return 0;
    } else {
        *inside = 0; /* shape.slang:200: unsupported shapes return the point itself */
        memcpy(out, pt, sizeof(float) * 3);
        return 1;
    }
    sim_mul_pt(pose, lproj, out);
    return 1;
}

int so_project_point_on_boundary(const so_shape *sh, const so_sim3 *pose, const float pt[3], const so_mesh *mesh,
                                 float max_dist, float out[3], int *inside) {
    *inside = 0;
    return shape_project_point_on_boundary(sh, pose, pt, mesh, max_dist, out, inside);
}

void so_set_collision_mesh(so_state *s, const float *vtx4, size_t num_vertices, const uint32_t *idx,
                           size_t num_indices) {
    free((void *)s->mesh.vtx);
    free((void *)s->mesh.idx);
    s->mesh.vtx = NULL;
    s->mesh.idx = NULL;
    if (!num_vertices || !num_indices) return;
    float *v = (float *)malloc(sizeof(float) * 4 * num_vertices);
    uint32_t *i = (uint32_t *)malloc(sizeof(uint32_t) * num_indices);
    memcpy(v, vtx4, sizeof(float) * 4 * num_vertices);
    memcpy(i, idx, sizeof(uint32_t) * num_indices);
    s->mesh.vtx = v;
    s->mesh.idx = i;
    s->mesh.nv = num_vertices;
    s->mesh.ni = num_indices;
}

/* collide.slang:87-154 (non-trimesh branch) */
static ncoll collide(const so_state *s, const float point[3]) {
    ncoll c;
    memset(&c, 0, sizeof c);
    float dist_cap = s->h * 1.5f;
    float best_dist = 1.0e10f;
    for (size_t i = 0; i < s->nb && i < 16; i++) {
        float proj[3];
        int inside = 0;
        if (!shape_project_point_on_boundary(&s->bodies[i].shape, &s->bodies[i].pose, point, &s->mesh, dist_cap, proj,
                                             &inside))
            continue;
        float dpt[3] = {proj[0] - point[0], proj[1] - point[1], proj[2] - point[2]};
        if (inside || (fabsf(dpt[0]) <= dist_cap && fabsf(dpt[1]) <= dist_cap && fabsf(dpt[2]) <= dist_cap)) {
            float sign = inside ? -1.0f : 1.0f;
            float dist = len3(dpt);
            if (sign * dist < best_dist && dist != 0.0f) {
                for (int k = 0; k < 3; k++) c.normal[k] = dpt[k] / (dist * -sign);
                memcpy(c.point, proj, sizeof proj);
                c.collider_id = (uint32_t)i;
                c.valid = 1;
                best_dist = dist;
            }
        }
    }
    return c;
}

/* boundary_condition.slang:31-89 */
static void project_velocity(so_bc bc, float vel[3], const float n[3], float dist, float dt) {
    switch (bc.ty) {
    case 0: /* Stick */
        if (dist <= 0.0f) vel[0] = vel[1] = vel[2] = 0.0f;
        return;
    case 1: /* Slip */
        if (dist <= 0.0f) {
            float nv = dot3(vel, n);
            for (int i = 0; i < 3; i++) vel[i] = vel[i] - n[i] * nv;
        }
        return;
    case 2: /* Separate */
    case 3: { /* FrictionZUp: same projection; the friction flag is computed but unused */
        float nv = dot3(vel, n);
        if (nv < 0.0f) {
            if (dist <= 0.0f) {
                float tv[3] = {vel[0] - n[0] * nv, vel[1] - n[1] * nv, vel[2] - n[2] * nv};
                float tl = len3(tv);
                float dir[3] = {0.0f, 0.0f, 0.0f};
                if (tl > 1.0e-8f)
                    for (int i = 0; i < 3; i++) dir[i] = tv[i] / tl;
                float m = fmaxf(0.0f, tl + bc.friction * nv);
                for (int i = 0; i < 3; i++) vel[i] = dir[i] * m;
            } else if (-nv * dt > dist) {
                float k = dist / dt + nv;
                for (int i = 0; i < 3; i++) vel[i] = vel[i] - k * n[i];
            }
        }
        return;
    }
    default:
        return;
    }
}

/* ------------------------------------------------------------------------- */
/* Grid update, shaders/slosh/solver/grid_update.slang:38-163                   */

static void grid_update(so_state *s) {
    const float h = s->h;
    const float dt = s->params.dt;
    const float *g = s->params.gravity;
#pragma omp parallel for schedule(static)
    for (long bid = 0; bid < (long)s->nblocks; bid++) {
        const header *hd = &s->hdr[bid];
        uint32_t prev_hid = find_block_header_id(s->prev_hmap, s->cap, hd->vid);
        for (uint32_t tid = 0; tid < 64u; tid++) {
            int lc[3] = {(int)(tid % 4u), (int)((tid / 4u) % 4u), (int)(tid / 16u)};
            float cell_pos[3];
            for (int i = 0; i < 3; i++) cell_pos[i] = (float)(hd->vid[i] * 4 + lc[i]) * h;
            size_t gid = (size_t)bid * 64u + tid;
            float *mvm = s->nodes[gid];
            /* update_single_cell, grid_update.slang:154-163 */
            float mass = mvm[3];
            float inv_mass = mass > 0.0f ? 1.0f / mass : 0.0f;
            float v[3];
            float lim = h / dt;
            for (int i = 0; i < 3; i++) {
                v[i] = (mvm[i] + mass * g[i] * dt) * inv_mass;
                v[i] = fminf(fmaxf(v[i], -lim), lim);
            }
            ncoll c;
            if (prev_hid != NONE) c = s->prev_coll[(size_t)prev_hid * 64u + tid];
            else c = collide(s, cell_pos);
            s->coll[gid] = c;
            if (c.valid) {
                so_bc bc = s->mats[c.collider_id];
                float d[3] = {cell_pos[0] - c.point[0], cell_pos[1] - c.point[1], cell_pos[2] - c.point[2]};
                float dist = dot3(d, c.normal) - h;
                project_velocity(bc, v, c.normal, dist, dt);
            }
            mvm[0] = v[0]; mvm[1] = v[1]; mvm[2] = v[2]; mvm[3] = mass;
        }
    }
}

/* Single-node grid update without the collision cache (known-answer tests). */
void so_grid_update_node(const so_body *bodies, const so_bc *mats, size_t nb, float h, const so_sim_params *params,
                         const float cell_pos[3], float mvm[4], float out_normal_dist[4]) {
    so_state tmp;
    memset(&tmp, 0, sizeof tmp);
    tmp.h = h;
    tmp.nb = nb;
    tmp.bodies = (so_body *)bodies;
    tmp.params = *params;
    const float dt = params->dt;
    float mass = mvm[3];
    float inv_mass = mass > 0.0f ? 1.0f / mass : 0.0f;
    float lim = h / dt;
    float v[3];
    for (int i = 0; i < 3; i++) {
        v[i] = (mvm[i] + mass * params->gravity[i] * dt) * inv_mass;
        v[i] = fminf(fmaxf(v[i], -lim), lim);
    }
    ncoll c = collide(&tmp, cell_pos);
    out_normal_dist[0] = out_normal_dist[1] = out_normal_dist[2] = 0.0f;
    out_normal_dist[3] = 1.0e10f;
    if (c.valid) {
        float d[3] = {cell_pos[0] - c.point[0], cell_pos[1] - c.point[1], cell_pos[2] - c.point[2]};
        float dist = dot3(d, c.normal) - h;
        project_velocity(mats[c.collider_id], v, c.normal, dist, dt);
        memcpy(out_normal_dist, c.normal, sizeof c.normal);
        out_normal_dist[3] = dist;
    }
    mvm[0] = v[0]; mvm[1] = v[1]; mvm[2] = v[2];
}

/* ------------------------------------------------------------------------- */
/* G2P, shaders/slosh/solver/g2p.slang:34-265                                   */

static void g2p(so_state *s) {
    const float cell_width = s->h;
#pragma omp parallel for schedule(dynamic, 4)
    for (long bid = 0; bid < (long)s->nblocks; bid++) {
        const header *hd = &s->hdr[bid];
        float shared[216][4];
        /* global_shared_memory_transfers, g2p.slang:78-148 */
        for (uint32_t tz = 0; tz < 4; tz++)
            for (uint32_t ty = 0; ty < 4; ty++)
                for (uint32_t tx = 0; tx < 4; tx++)
                    for (uint32_t i = 0; i <= 1; i++)
                        for (uint32_t j = 0; j <= 1; j++)
                            for (uint32_t k = 0; k <= 1; k++) {
                                if ((i == 1 && tx > 1) || (j == 1 && ty > 1) || (k == 1 && tz > 1)) continue;
                                int32_t ob[3] = {hd->vid[0] + (int32_t)i, hd->vid[1] + (int32_t)j, hd->vid[2] + (int32_t)k};
                                uint32_t ohid = find_block_header_id(s->hmap, s->cap, ob);
                                uint32_t si = (i * 4 + tx) + (j * 4 + ty) * 6 + (k * 4 + tz) * 36;
                                if (ohid != NONE) {
                                    memcpy(shared[si], s->nodes[(size_t)ohid * 64u + tx + ty * 4 + tz * 16], sizeof(float) * 4);
                                } else {
                                    shared[si][0] = shared[si][1] = shared[si][2] = shared[si][3] = 0.0f;
                                }
                            }
        uint32_t end = hd->first_particle + hd->num_particles;
        for (uint32_t sp = hd->first_particle; sp < end; sp++) {
            uint32_t pid = s->sorted[sp];
            so_kinematics *kin = &s->kin[pid];
            /* particle_g2p, g2p.slang:150-265 */
            float mvm[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            m3 vg = m3_zero();
            float vgd = 0.0f;
            if (kin->enabled != 0) {
                const float *pos = s->pos[pid];
                float inv_d = 4.0f / (cell_width * cell_width);
                float ref[3], w[3][3];
                uint32_t assoc[3];
                for (int i = 0; i < 3; i++) {
                    ref[i] = (rintf(pos[i] / cell_width) - 1.0f) * cell_width - pos[i];
                    k_eval_all(-ref[i] / cell_width, w[i]);
                }
                associated_cell_index_in_block_off_by_one(cell_width, pos, assoc);
                uint32_t packed = assoc[0] + assoc[1] * 6 + assoc[2] * 36;
                for (int i = 0; i < 27; i++) {
                    const uint32_t *sh = NBH_SHIFTS[i];
                    uint32_t sid = packed + sh[0] + sh[1] * 6 + sh[2] * 36;
                    const float *cd = shared[sid];
                    float dpt[3];
                    for (int k = 0; k < 3; k++) dpt[k] = ref[k] + (float)sh[k] * cell_width;
                    float weight = w[0][sh[0]] * w[1][sh[1]] * w[2][sh[2]];
                    for (int k = 0; k < 4; k++) mvm[k] += cd[k] * weight;
                    float wi = weight * inv_d;
                    /* outer_product(a, b): row r = a * b[r] */
                    for (int r = 0; r < 3; r++)
                        for (int c2 = 0; c2 < 3; c2++) vg.m[r][c2] += wi * (cd[c2] * dpt[r]);
                    vgd += weight * inv_d * dot3(cd, dpt);
                }
            }
            m3_store(&kin->affine, vg);
            kin->vel_grad_det = vgd;
            kin->velocity[0] = mvm[0];
            kin->velocity[1] = mvm[1];
            kin->velocity[2] = mvm[2];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* 3x3 SVD (stand-in for stensor's Svd3, which is not vendored).                 */
/* Cyclic Jacobi on A^T A, then Givens QR of A V (McAdams et al. 2011 layout):   */
/* U, V are rotations, sigma_0 >= sigma_1 >= |sigma_2|, sign(sigma_2) = sign(det A). */
/* The sweeps stop once the off-diagonal part is below TOL times the diagonal     */
/* (squared norms): 1e-14 in FP32, 1e-30 in FP64, i.e. converged to round-off.   */

#define SVD_DEF(T, SQRT, FABS, NAME, TOL)                                                       \
    static void NAME(const T a[9], T u[9], T sig[3], T v[9]) {                                   \
        T s[3][3];                                                                               \
        for (int i = 0; i < 3; i++)                                                              \
            for (int j = 0; j < 3; j++)                                                          \
                s[i][j] = a[0 * 3 + i] * a[0 * 3 + j] + a[1 * 3 + i] * a[1 * 3 + j] + a[2 * 3 + i] * a[2 * 3 + j]; \
        T vv[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};                                          \
        static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};                                    \
        for (int sweep = 0; sweep < 12; sweep++) {                                               \
            T off = s[0][1] * s[0][1] + s[0][2] * s[0][2] + s[1][2] * s[1][2];                   \
            T dg = s[0][0] * s[0][0] + s[1][1] * s[1][1] + s[2][2] * s[2][2];                    \
            if (!(off > (T)(TOL) * dg)) break;                                                   \
            for (int e = 0; e < 3; e++) {                                                        \
                int p = PQ[e][0], q = PQ[e][1];                                                  \
                T apq = s[p][q];                                                                 \
                if (apq == (T)0) continue;                                                       \
                T theta = (s[q][q] - s[p][p]) / ((T)2 * apq);                                    \
                T t = (theta >= (T)0 ? (T)1 : (T)-1) / (FABS(theta) + SQRT(theta * theta + (T)1)); \
                T c = (T)1 / SQRT(t * t + (T)1);                                                 \
                T sn = t * c;                                                                    \
                for (int k = 0; k < 3; k++) {                                                    \
                    T skp = s[k][p], skq = s[k][q];                                              \
                    s[k][p] = c * skp - sn * skq;                                                \
                    s[k][q] = sn * skp + c * skq;                                                \
                }                                                                                \
                for (int k = 0; k < 3; k++) {                                                    \
                    T spk = s[p][k], sqk = s[q][k];                                              \
                    s[p][k] = c * spk - sn * sqk;                                                \
                    s[q][k] = sn * spk + c * sqk;                                                \
                }                                                                                \
                for (int k = 0; k < 3; k++) {                                                    \
                    T vkp = vv[k][p], vkq = vv[k][q];                                            \
                    vv[k][p] = c * vkp - sn * vkq;                                               \
                    vv[k][q] = sn * vkp + c * vkq;                                               \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
        /* sort eigenvalues descending (swap columns, keep det V = +1 by negating) */            \
        T ev[3] = {s[0][0], s[1][1], s[2][2]};                                                   \
        for (int pass = 0; pass < 2; pass++)                                                     \
            for (int i = 0; i < 2; i++)                                                          \
                if (ev[i] < ev[i + 1]) {                                                         \
                    T te = ev[i]; ev[i] = ev[i + 1]; ev[i + 1] = te;                             \
                    for (int k = 0; k < 3; k++) {                                                \
                        T tv = vv[k][i]; vv[k][i] = -vv[k][i + 1]; vv[k][i + 1] = tv;            \
                    }                                                                            \
                }                                                                                \
        /* B = A V */                                                                            \
        T b[3][3];                                                                               \
        for (int i = 0; i < 3; i++)                                                              \
            for (int j = 0; j < 3; j++)                                                          \
                b[i][j] = a[i * 3 + 0] * vv[0][j] + a[i * 3 + 1] * vv[1][j] + a[i * 3 + 2] * vv[2][j]; \
        /* Givens QR: zero b[1][0], b[2][0], b[2][1]; Q accumulates as a rotation. */            \
        T qm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};                                          \
        static const int GR[3][2] = {{0, 1}, {0, 2}, {1, 2}};                                    \
        for (int e = 0; e < 3; e++) {                                                            \
            int r0 = GR[e][0], r1 = GR[e][1], col = GR[e][0];                                    \
            T x = b[r0][col], y = b[r1][col];                                                    \
            T r = SQRT(x * x + y * y);                                                           \
            T c = (T)1, sn = (T)0;                                                               \
            if (r > (T)0) { c = x / r; sn = y / r; }                                             \
            for (int k = 0; k < 3; k++) {                                                        \
                T b0 = b[r0][k], b1 = b[r1][k];                                                  \
                b[r0][k] = c * b0 + sn * b1;                                                     \
                b[r1][k] = -sn * b0 + c * b1;                                                    \
            }                                                                                    \
            for (int k = 0; k < 3; k++) {                                                        \
                T q0 = qm[k][r0], q1 = qm[k][r1];                                                \
                qm[k][r0] = c * q0 + sn * q1;                                                    \
                qm[k][r1] = -sn * q0 + c * q1;                                                   \
            }                                                                                    \
        }                                                                                        \
        for (int i = 0; i < 3; i++) {                                                            \
            sig[i] = b[i][i];                                                                    \
            for (int j = 0; j < 3; j++) {                                                        \
                u[i * 3 + j] = qm[i][j];                                                         \
                v[i * 3 + j] = vv[i][j];                                                         \
            }                                                                                    \
        }                                                                                        \
    }

SVD_DEF(float, sqrtf, fabsf, svd3_f32, 1.0e-14)
SVD_DEF(double, sqrt, fabs, svd3_f64, 1.0e-30)

void so_svd3(const double a[9], int f64, double u[9], double sig[3], double v[9]) {
    if (f64) {
        svd3_f64(a, u, sig, v);
    } else {
        float af[9], uf[9], sf[3], vf[9];
        for (int i = 0; i < 9; i++) af[i] = (float)a[i];
        svd3_f32(af, uf, sf, vf);
        for (int i = 0; i < 9; i++) { u[i] = uf[i]; v[i] = vf[i]; }
        for (int i = 0; i < 3; i++) sig[i] = sf[i];
    }
}

typedef struct { m3 u, vt; float s[3]; } svd_t;

static svd_t svd_of(m3 f, int svd_mode) {
    svd_t r;
    if (svd_mode == 1) {
        double a[9], u[9], sg[3], v[9];
        for (int i = 0; i < 9; i++) a[i] = f.m[i / 3][i % 3];
        svd3_f64(a, u, sg, v);
        for (int i = 0; i < 3; i++) {
            r.s[i] = (float)sg[i];
            for (int j = 0; j < 3; j++) { r.u.m[i][j] = (float)u[i * 3 + j]; r.vt.m[j][i] = (float)v[i * 3 + j]; }
        }
    } else {
        float a[9], u[9], sg[3], v[9];
        for (int i = 0; i < 9; i++) a[i] = f.m[i / 3][i % 3];
        svd3_f32(a, u, sg, v);
        for (int i = 0; i < 3; i++) {
            r.s[i] = sg[i];
            for (int j = 0; j < 3; j++) { r.u.m[i][j] = u[i * 3 + j]; r.vt.m[j][i] = v[i * 3 + j]; }
        }
    }
    return r;
}

/* Svd3::recompose: U diag(s) Vt */
static m3 svd_recompose(const m3 *u, const float s[3], const m3 *vt) {
    m3 us;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) us.m[i][j] = u->m[i][j] * s[j];
    return m3_mul(us, *vt);
}

/* ------------------------------------------------------------------------- */
/* Material models, shaders/slosh/models                                        */

static float u2f(uint32_t x) { float f; memcpy(&f, &x, 4); return f; }
static uint32_t f2u(float f) { uint32_t x; memcpy(&x, &f, 4); return x; }

/* neo_hookean_elasticity.slang:16-27 */
static m3 nh_stress(float lambda, float mu, m3 f) {
    float j = fmaxf(m3_det(f), 1.0e-10f);
    float diag = lambda * logf(j) - mu;
    m3 st = m3_scale(m3_mul(m3_t(f), f), mu);
    st.m[0][0] += diag; st.m[1][1] += diag; st.m[2][2] += diag;
    return st;
}

/* linear_elasticity.slang:135-148 (fixed-corotated) */
static m3 fc_stress(float lambda, float mu, m3 f, int svd_mode) {
    svd_t svd = svd_of(f, svd_mode);
    float j = svd.s[0] * svd.s[1] * svd.s[2];
    float sm1[3] = {svd.s[0] - 1.0f, svd.s[1] - 1.0f, svd.s[2] - 1.0f};
    float diag = lambda * (j - 1.0f) * j;
    m3 r = m3_scale(m3_mul(m3_t(f), svd_recompose(&svd.u, sm1, &svd.vt)), 2.0f * mu);
    r.m[0][0] += diag; r.m[1][1] += diag; r.m[2][2] += diag;
    return r;
}

/* drucker_prager.slang:46-50 */
static float dp_alpha(const float *hp, float q) {
    float angle = hp[0] + (hp[1] * q - hp[3]) * expf(-hp[2] * q);
    float s_angle = sinf(angle);
    return sqrtf(2.0f / 3.0f) * (2.0f * s_angle) / (3.0f - s_angle);
}

/* drucker_prager.slang:144-200. state = {pdet, hardening, log_vol_gain};
   dp = {h0,h1,h2,h3,lambda,mu}. Returns the projected deformation gradient. */
static m3 dp_project(float state[3], const float dp[6], m3 f, int svd_mode) {
    float lambda = dp[4], mu = dp[5];
    if (lambda == 0.0f) return f;
    svd_t svd = svd_of(f, svd_mode);
    float alpha = dp_alpha(dp, state[1]);
    /* project_deformation_gradient, drucker_prager.slang:144-163 */
    const float d = 3.0f;
    float strain[3], dev[3];
    float lvg = state[2];
    for (int i = 0; i < 3; i++) strain[i] = logf(svd.s[i]) + lvg / d;
    float tr = strain[0] + strain[1] + strain[2];
    for (int i = 0; i < 3; i++) dev[i] = strain[i] - tr / d;
    float new_s[3];
    float hard;
    int valid;
    if (tr > 0.0f || (dev[0] == 0.0f && dev[1] == 0.0f && dev[2] == 0.0f)) {
        new_s[0] = new_s[1] = new_s[2] = 1.0f;
        hard = len3(strain);
        valid = 1;
    } else {
        float dn = len3(dev);
        float gamma = dn + (d * lambda + 2.0f * mu) / (2.0f * mu) * tr * alpha;
        if (gamma <= 0.0f) {
            valid = 0;
            hard = 0.0f;
        } else {
            for (int i = 0; i < 3; i++) new_s[i] = expf(strain[i] - dev[i] * (gamma / dn));
            hard = gamma;
            valid = 1;
        }
    }
    if (!valid) return f;
    float prev_det = svd.s[0] * svd.s[1] * svd.s[2];
    float new_det = new_s[0] * new_s[1] * new_s[2];
    state[0] = state[0] * prev_det / new_det;
    state[2] = state[2] + logf(prev_det) - logf(new_det);
    state[1] = state[1] + hard;
    return svd_recompose(&svd.u, new_s, &svd.vt);
}

/* FP64 variants of the SVD-dependent steps (svd_mode 1): the SVD, F - R and the
   Drucker-Prager return mapping are evaluated in double and rounded once. */
typedef struct { double m[3][3]; } d3;

static d3 d3_from(m3 f) { d3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = f.m[i][j]; return r; }
static m3 m3_from(d3 f) { m3 r; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = (float)f.m[i][j]; return r; }

static void svd_d(d3 f, d3 *u, double s[3], d3 *v) {
    double a[9], uu[9], vv[9];
    for (int i = 0; i < 9; i++) a[i] = f.m[i / 3][i % 3];
    svd3_f64(a, uu, s, vv);
    for (int i = 0; i < 9; i++) { u->m[i / 3][i % 3] = uu[i]; v->m[i / 3][i % 3] = vv[i]; }
}

static d3 recompose_d(const d3 *u, const double *s, const d3 *v) {
    d3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.m[i][j] = (u->m[i][0] * s[0]) * v->m[j][0] + (u->m[i][1] * s[1]) * v->m[j][1] +
                        (u->m[i][2] * s[2]) * v->m[j][2];
    return r;
}

static m3 fc_stress_f64(float lambda, float mu, m3 f32) {
    d3 f = d3_from(f32);
    d3 u, v;
    double s[3];
    svd_d(f, &u, s, &v);
    double j = s[0] * s[1] * s[2];
    double sm1[3] = {s[0] - 1.0, s[1] - 1.0, s[2] - 1.0};
    d3 rc = recompose_d(&u, sm1, &v);
    d3 r;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            r.m[i][k] = 2.0 * (double)mu * (f.m[0][i] * rc.m[0][k] + f.m[1][i] * rc.m[1][k] + f.m[2][i] * rc.m[2][k]);
    double diag = (double)lambda * (j - 1.0) * j;
    r.m[0][0] += diag; r.m[1][1] += diag; r.m[2][2] += diag;
    return m3_from(r);
}

static m3 dp_project_f64(float state[3], const float dp[6], m3 f32) {
    double lambda = dp[4], mu = dp[5];
    if (dp[4] == 0.0f) return f32;
    d3 f = d3_from(f32);
    d3 u, v;
    double s[3];
    svd_d(f, &u, s, &v);
    double q = state[1];
    double angle = (double)dp[0] + ((double)dp[1] * q - (double)dp[3]) * exp(-(double)dp[2] * q);
    double sa = sin(angle);
    double alpha = sqrt(2.0 / 3.0) * (2.0 * sa) / (3.0 - sa);
    const double d = 3.0;
    double strain[3], dev[3], ns[3], hard;
    for (int i = 0; i < 3; i++) strain[i] = log(s[i]) + (double)state[2] / d;
    double tr = strain[0] + strain[1] + strain[2];
    for (int i = 0; i < 3; i++) dev[i] = strain[i] - tr / d;
    if (tr > 0.0 || (dev[0] == 0.0 && dev[1] == 0.0 && dev[2] == 0.0)) {
        ns[0] = ns[1] = ns[2] = 1.0;
        hard = sqrt(strain[0] * strain[0] + strain[1] * strain[1] + strain[2] * strain[2]);
    } else {
        double dn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
        double gamma = dn + (d * lambda + 2.0 * mu) / (2.0 * mu) * tr * alpha;
        if (gamma <= 0.0) return f32;
        for (int i = 0; i < 3; i++) ns[i] = exp(strain[i] - dev[i] * (gamma / dn));
        hard = gamma;
    }
    double prev_det = s[0] * s[1] * s[2];
    double new_det = ns[0] * ns[1] * ns[2];
    state[0] = (float)((double)state[0] * prev_det / new_det);
    state[2] = (float)((double)state[2] + log(prev_det) - log(new_det));
    state[1] = (float)((double)state[1] + hard);
    return m3_from(recompose_d(&u, ns, &v));
}

/* DefaultParticleModel::update, default.slang:51-92 */
static m3 model_update(so_model *model, m3 *f, int svd_mode) {
    uint32_t tag = model->w[0];
    float p[12];
    for (int i = 0; i < 12; i++) p[i] = u2f(model->w[1 + i]);
    switch (tag) {
    case 0:
        return svd_mode == 1 ? fc_stress_f64(p[0], p[1], *f) : fc_stress(p[0], p[1], *f, svd_mode);
    case 1:
        return nh_stress(p[0], p[1], *f);
    case 2:
    case 3: {
        float state[3] = {p[0], p[1], p[2]};
        m3 nf = svd_mode == 1 ? dp_project_f64(state, &p[3], *f) : dp_project(state, &p[3], *f, svd_mode);
        for (int i = 0; i < 3; i++) model->w[1 + i] = f2u(state[i]);
        m3 st = tag == 2 ? (svd_mode == 1 ? fc_stress_f64(p[9], p[10], nf) : fc_stress(p[9], p[10], nf, svd_mode))
                         : nh_stress(p[9], p[10], nf);
        *f = nf;
        return st;
    }
    default:
        return m3_zero();
    }
}

void so_model_update(so_model *model, float def_grad[9], float stress[9], int svd_mode) {
    m3 f;
    for (int i = 0; i < 9; i++) f.m[i / 3][i % 3] = def_grad[i];
    m3 st = model_update(model, &f, svd_mode);
    for (int i = 0; i < 9; i++) {
        def_grad[i] = f.m[i / 3][i % 3];
        stress[i] = st.m[i / 3][i % 3];
    }
}

/* ------------------------------------------------------------------------- */
/* Particle update, shaders/slosh/solver/particle_update.slang:32-160           */

static void particle_update(so_state *s) {
    const float dt = s->params.dt;
    const float cell_width = s->h;
#pragma omp parallel for schedule(dynamic, 256)
    for (long id = 0; id < (long)s->n; id++) {
        so_kinematics *kin = &s->kin[id];
        m3 def_grad = m3_load(&s->F[id]);
        const so_props *props = &s->props[id];
        float *pp = s->pos[id];
        float v[3] = {kin->velocity[0], kin->velocity[1], kin->velocity[2]};
        float vl = len3(v);
        if (vl > cell_width / dt) {
            float l2 = len3(v);
            for (int i = 0; i < 3; i++) v[i] = v[i] / l2 * cell_width / dt;
        }
        for (int i = 0; i < 3; i++) v[i] = v[i] / (1.0f + props->damping * dt);
        if (props->fixed != 0) v[0] = v[1] = v[2] = 0.0f;
        float new_pos[3];
        for (int i = 0; i < 3; i++) new_pos[i] = pp[i] + v[i] * dt;
        m3 aff = m3_load(&kin->affine);
        /* FLUID is never set (default.slang:47-49). */
        def_grad = m3_add(def_grad, m3_mul(def_grad, m3_scale(aff, dt)));
        if (props->fixed != 0) aff = m3_zero();
        m3 stress = model_update(&s->models[id], &def_grad, s->svd_mode);
        float inv_d = 4.0f / (cell_width * cell_width);
        m3 affine = m3_sub(m3_scale(aff, kin->mass), m3_scale(stress, props->init_volume * inv_d * dt));
        int ok = new_pos[0] == new_pos[0] && new_pos[1] == new_pos[1] && new_pos[2] == new_pos[2] &&
                 m3_det(def_grad) > 0.0f;
        if (ok) {
            pp[0] = new_pos[0]; pp[1] = new_pos[1]; pp[2] = new_pos[2];
            aff = affine;
        } else {
            kin->enabled = 0;
            v[0] = v[1] = v[2] = 0.0f;
            def_grad = m3_ident();
            aff = m3_zero();
            kin->mass = 0.0f;
        }
        kin->force_dt[0] = kin->force_dt[1] = kin->force_dt[2] = 0.0f;
        kin->velocity[0] = v[0]; kin->velocity[1] = v[1]; kin->velocity[2] = v[2];
        m3_store(&kin->affine, aff);
        m3_store(&s->F[id], def_grad);
    }
}

/* ------------------------------------------------------------------------- */
/* Rigid-body integration, shaders/slosh/solver/rigid_impulses.slang:70-121     */

static void quat_mul(const float a[4], const float b[4], float out[4]) {
    out[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
    out[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
    out[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
    out[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}

static void rigid_update(so_state *s) {
    const float dt = s->params.dt;
    for (size_t id = 0; id < s->nb && id < 16; id++) {
        so_body *b = &s->bodies[id];
        float lin[3], ang[3];
        for (int i = 0; i < 3; i++) {
            lin[i] = (float)s->imp[id][i] / 1.0e5f; /* int2flt */
            ang[i] = (float)s->imp[id][3 + i] / 1.0e5f;
        }
        memset(s->imp[id], 0, sizeof s->imp[id]);
        /* apply_impulse, body.slang:63-71 */
        float nl[3], na[3];
        for (int i = 0; i < 3; i++) {
            nl[i] = b->vel.linear[i] + b->mprops.inv_mass[i] * lin[i];
            float aa = ang[0] * b->mprops.inv_inertia[0][i] + ang[1] * b->mprops.inv_inertia[1][i] +
                       ang[2] * b->mprops.inv_inertia[2][i];
            na[i] = b->vel.angular[i] + aa;
        }
        float ln = len3(nl), an = len3(na);
        float lin_limit = 0.1f * s->h / dt;
        float ang_limit = 1.0f;
        if (len3(lin) != 0.0f || len3(ang) != 0.0f) {
            if (ln > lin_limit)
                for (int i = 0; i < 3; i++) nl[i] = nl[i] * (lin_limit / ln);
            if (an > ang_limit)
                for (int i = 0; i < 3; i++) na[i] = na[i] * (ang_limit / an);
        }
        /* integrate_velocity, body.slang:117-130 */
        float init_com[3];
        sim_mul_pt(&b->pose, b->local_mprops.com, init_com);
        float init_scale = b->pose.translation_scale[3];
        float sa[3] = {na[0] * dt, na[1] * dt, na[2] * dt};
        float angle = len3(sa);
        float dq[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        if (angle != 0.0f) {
            float sh = sinf(angle * 0.5f) / angle;
            dq[0] = sa[0] * sh; dq[1] = sa[1] * sh; dq[2] = sa[2] * sh;
            dq[3] = cosf(angle * 0.5f);
        }
        float rel[3] = {b->pose.translation_scale[0] - init_com[0], b->pose.translation_scale[1] - init_com[1],
                        b->pose.translation_scale[2] - init_com[2]};
        float rr[3];
        quat_rotate(dq, rel, rr);
        float nq[4];
        quat_mul(dq, b->pose.rotation, nq);
        float sq = nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3];
        float k = 0.5f * (3.0f - sq); /* renormalize_fast */
        for (int i = 0; i < 4; i++) b->pose.rotation[i] = nq[i] * k;
        for (int i = 0; i < 3; i++) b->pose.translation_scale[i] = init_com[i] + rr[i] * init_scale + nl[i] * dt;
        /* gravity, masked per component by inv_mass != 0 */
        for (int i = 0; i < 3; i++) {
            float mask = b->mprops.inv_mass[i] != 0.0f ? 1.0f : 0.0f;
            nl[i] += s->params.gravity[i] * mask * dt;
        }
        memcpy(b->vel.linear, nl, sizeof nl);
        memcpy(b->vel.angular, na, sizeof na);
    }
}

/* ------------------------------------------------------------------------- */
/* Pipeline: MpmPipeline::launch_step order, src/pipeline.rs:562-770            */

int so_step(so_state *s) {
    s->overflow = 0;
    grid_sort(s);
    p2g(s);
    grid_update(s);
    g2p(s);
    particle_update(s);
    if (s->nb > 0) rigid_update(s);
    return s->overflow ? -1 : 0;
}

so_state *so_create(const so_sim_params *params, size_t n, const float *pos4, const so_kinematics *kin,
                    const so_mat3 *def_grad, const so_props *props, const so_model *models, size_t nbodies,
                    const so_body *bodies, const so_bc *materials, float cell_width, uint32_t grid_capacity,
                    int svd_mode) {
    so_state *s = (so_state *)calloc(1, sizeof(so_state));
    if (!s) return NULL;
    uint32_t cap = 1;
    while (cap < grid_capacity) cap <<= 1; /* next_power_of_two, grid.rs:437 */
    s->n = n;
    s->h = cell_width;
    s->cap = cap;
    s->params = *params;
    s->svd_mode = svd_mode;
    size_t nn = n ? n : 1;
    s->pos = malloc(sizeof(float) * 4 * nn);
    s->kin = malloc(sizeof(so_kinematics) * nn);
    s->F = malloc(sizeof(so_mat3) * nn);
    s->props = malloc(sizeof(so_props) * nn);
    s->models = malloc(sizeof(so_model) * nn);
    if (n) {
        memcpy(s->pos, pos4, sizeof(float) * 4 * n);
        memcpy(s->kin, kin, sizeof(so_kinematics) * n);
        memcpy(s->F, def_grad, sizeof(so_mat3) * n);
        memcpy(s->props, props, sizeof(so_props) * n);
        memcpy(s->models, models, sizeof(so_model) * n);
    }
    s->nb = nbodies;
    s->bodies = malloc(sizeof(so_body) * (nbodies ? nbodies : 1));
    s->mats = malloc(sizeof(so_bc) * (nbodies ? nbodies : 1));
    if (nbodies) {
        memcpy(s->bodies, bodies, sizeof(so_body) * nbodies);
        memcpy(s->mats, materials, sizeof(so_bc) * nbodies);
    }
    s->hmap = malloc(sizeof(hentry) * cap);
    s->prev_hmap = malloc(sizeof(hentry) * cap);
    for (uint32_t i = 0; i < cap; i++) {
        s->hmap[i].state = NONE;
        s->prev_hmap[i].state = NONE;
    }
    s->hdr = calloc(cap, sizeof(header));
    s->scan = calloc(cap, sizeof(uint32_t));
    s->sorted = calloc(nn * 8u, sizeof(uint32_t));
    s->nodes = calloc((size_t)cap * 64u, sizeof(float) * 4);
    s->coll = calloc((size_t)cap * 64u, sizeof(ncoll));
    s->prev_coll = calloc((size_t)cap * 64u, sizeof(ncoll));
    s->pblk = malloc(sizeof(*s->pblk) * nn);
    s->passoc = malloc(sizeof(*s->passoc) * nn);
    s->phid = malloc(sizeof(uint32_t) * nn);
    return s;
}

void so_destroy(so_state *s) {
    if (!s) return;
    free(s->pos); free(s->kin); free(s->F); free(s->props); free(s->models);
    free(s->bodies); free(s->mats); free(s->hmap); free(s->prev_hmap); free(s->hdr);
    free(s->scan); free(s->sorted); free(s->nodes); free(s->coll); free(s->prev_coll);
    free((void *)s->mesh.vtx); free((void *)s->mesh.idx);
    free(s->pblk); free(s->passoc); free(s->phid);
    free(s);
}

void so_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

void so_read_positions(const so_state *s, float *out4) { memcpy(out4, s->pos, sizeof(float) * 4 * s->n); }
void so_read_kinematics(const so_state *s, so_kinematics *out) { memcpy(out, s->kin, sizeof(so_kinematics) * s->n); }
void so_read_def_grad(const so_state *s, so_mat3 *out) { memcpy(out, s->F, sizeof(so_mat3) * s->n); }
void so_read_models(const so_state *s, so_model *out) { memcpy(out, s->models, sizeof(so_model) * s->n); }
void so_read_bodies(const so_state *s, so_sim3 *poses, so_velocity *vels) {
    for (size_t i = 0; i < s->nb; i++) {
        if (poses) poses[i] = s->bodies[i].pose;
        if (vels) vels[i] = s->bodies[i].vel;
    }
}
/* ------------------------------------------------------------------------- */
/* Timestep bound, WgTimestepBounds (src/solver/timestep_bound.rs:56-102) and     */
/* shaders/slosh/solver/timestep_bound.slang:18-92 with the models' sound-speed   */
/* bound (models/utils.slang:100-136; default.slang:94-127).                      */

/* GpuTimestepBounds::secs_to_int (timestep_bound.slang:21-24), saturating. */
static uint32_t secs_to_int(float secs) {
    float x = floorf(secs * 1.0e12f);
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967295.0f) return 0xffffffffu;
    return (uint32_t)x;
}

uint32_t so_particle_timestep_bound(const so_model *model, const so_kinematics *kin, const so_mat3 *def_grad,
                                    const so_props *props, float cell_width) {
    float density0 = kin->mass / props->init_volume;
    m3 f = m3_load(def_grad);
    float det = m3_det(f);
    uint32_t tag = model->w[0];
    float dt;
    if (tag < 4u) {
        /* elastic part: (lambda, mu, cfl) at words 1..3 (elastic) or 10..12 (sand) */
        const uint32_t *e = model->w + (tag < 2u ? 1 : 10);
        float lambda, mu, cfl;
        memcpy(&lambda, e + 0, 4);
        memcpy(&mu, e + 1, 4);
        memcpy(&cfl, e + 2, 4);
        float bulk = lambda + 2.0f * mu / 3.0f; /* bulk_modulus_from_lame, utils.slang:26-28 */
        float shear = mu;                       /* shear_modulus_from_lame, utils.slang:22-24 */
        float curr_density = density0 / fmaxf(det, 1.0e-6f);
        float sound_speed = sqrtf((bulk + shear * 4.0f / 3.0f) / curr_density);
        float max_speed = fmaxf(len3(kin->velocity), sound_speed);
        dt = cfl * cell_width / max_speed;
    } else {
        dt = 0.0f; /* default.slang:124-125 */
    }
    dt = fminf(dt, cell_width / len3(kin->velocity));
    return secs_to_int(dt);
}

uint32_t so_timestep_bound(const so_state *s) {
    uint32_t best = 0xffffffffu; /* reset_timestep_bound */
    for (size_t id = 0; id < s->n; id++) {
        if (s->kin[id].enabled == 0) continue;
        uint32_t c = so_particle_timestep_bound(&s->models[id], &s->kin[id], &s->F[id], &s->props[id], s->h);
        if (c < best) best = c;
    }
    return best;
}

uint32_t so_num_blocks(const so_state *s) { return s->nblocks; }
void so_read_blocks(const so_state *s, so_block_info *out) {
    for (uint32_t i = 0; i < s->nblocks; i++) {
        const header *hd = &s->hdr[i];
        memcpy(out[i].vid, hd->vid, sizeof hd->vid);
        out[i].first_particle = hd->first_particle;
        out[i].num_particles = hd->num_particles;
        out[i].num_particles_with_extras = hd->num_particles_with_extras;
        memcpy(out[i].bucket_counts, hd->counts, sizeof hd->counts);
    }
}
void so_read_nodes(const so_state *s, float *out4) {
    memcpy(out4, s->nodes, sizeof(float) * 4 * (size_t)s->nblocks * 64u);
}
