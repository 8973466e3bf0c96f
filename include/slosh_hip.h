/*
 * slosh_hip.h -- C ABI of the MI355X (gfx950) MPM substep library libslosh_hip.so.
 *
 * Drop-in boundary for slosh's substep path. Each entry point names the
 * reference interface it replaces (paths relative to the slosh repository).
 * Plain pointers and sizes only; device streams are passed as the opaque
 * `slosh_stream` (a hipStream_t). Every function returns 0 on success or a
 * negative SLOSH_E_* code; `slosh_last_error()` gives the thread's last message.
 *
 * Record layouts are the reference's host/GPU layouts, byte for byte:
 *   positions  : vec4 per particle               (ParticlePosition, src/solver/particle.rs:388-389)
 *   kinematics : slosh_kinematics, 96 B          (Kinematics, src/solver/particle.rs:132-147)
 *   def_grad   : slosh_matrix, 3 x vec4 (48 B)   (Matrix; row r of the shader matrix = glam column r)
 *   properties : slosh_particle_properties, 20 B (ParticleProperties, src/solver/particle.rs:154-167)
 *   models     : slosh_particle_model, 52 B      (GpuParticleModel, src/solver/particle_model.rs:87-102)
 */
#ifndef SLOSH_HIP_H
#define SLOSH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLOSH_OK 0
#define SLOSH_E_INVALID -1       /* bad argument (null pointer, size, >16 colliders, ...) */
#define SLOSH_E_HIP -2           /* a HIP runtime call failed */
#define SLOSH_E_NO_MEMORY -3     /* device allocation failed */
#define SLOSH_E_GRID_OVERFLOW -4 /* sparse-grid hash map full (reference: silent, grid.slang:126-128) */
#define SLOSH_E_HOOK -5          /* a hook returned an error */
#define SLOSH_E_UNSUPPORTED -6   /* input the library does not implement (e.g. trimesh colliders) */
#define SLOSH_E_NO_KERNELS -7    /* pipeline built without builtin transfers and no hook handled the pass */
#define SLOSH_E_KEY_RANGE -8     /* a block coordinate outside the block-key space (reference: silent wrap) */

/* Material tags of GpuParticleModel (src/solver/particle_model.rs:89-100). */
#define SLOSH_MODEL_ELASTIC_LINEAR 0u
#define SLOSH_MODEL_ELASTIC_NEO_HOOKEAN 1u
#define SLOSH_MODEL_SAND_LINEAR 2u
#define SLOSH_MODEL_SAND_NEO_HOOKEAN 3u

/* Boundary-condition types (GpuBoundaryCondition, src/solver/boundary_condition.rs:12-17). */
#define SLOSH_BC_STICK 0u
#define SLOSH_BC_SLIP 1u
#define SLOSH_BC_SEPARATE 2u
#define SLOSH_BC_FRICTION_Z_UP 3u
#define SLOSH_BC_NON_REFLECTING 4u
#define SLOSH_BC_DISABLED 5u

/* Shape tags of GpuShape (src/rbd/shapes.rs:17-33). */
#define SLOSH_SHAPE_BALL 0u
#define SLOSH_SHAPE_CUBOID 1u

#define SLOSH_MAX_COLLIDERS 16 /* GpuMaterials::new assert, src/solver/boundary_condition.rs:77-80 */
#define SLOSH_MAX_SCENES 16383 /* batched key space: 14-bit scene id */
/* Block-key range. Single scene (pack_key, grid.slang:83-90): block coordinates
 * x, z in [-1023, 1023], y in [-511, 511]. Batched: 6 bits per axis, every block
 * coordinate in [-16, 46] (with dx = 1/128, positions in about [-0.5, 1.49]).
 * The "+1" neighbour of every block then has an exact key too. Creation and
 * append reject particles outside it (SLOSH_E_KEY_RANGE); a particle that moves
 * out of it is reported by slosh_check_errors (SLOSH_E_KEY_RANGE). */
#define SLOSH_BATCHED_BLOCK_MIN (-16)
#define SLOSH_BATCHED_BLOCK_MAX 46

typedef void *slosh_stream; /* hipStream_t; NULL = the default stream */
typedef struct slosh_pipeline slosh_pipeline;
typedef struct slosh_data slosh_data;
typedef struct slosh_timestamps slosh_timestamps;

/* MpmPipelineKernels, src/pipeline.rs:39-67. Non-zero = enabled. `cdf` and
 * `rigid_particles` are accepted and ignored: launch_step never dispatches them
 * (src/pipeline.rs:617-646). */
typedef struct {
    uint32_t builtin_transfers;
    uint32_t cdf;
    uint32_t rigid_particles;
    uint32_t node_reset;
    uint32_t grid_update;
    uint32_t particles_update;
} slosh_kernels;

/* SimulationParams, src/solver/params.rs:13-21 (3D). */
typedef struct {
    float gravity[3];
    float dt;
} slosh_sim_params;

typedef struct { float m[3][4]; } slosh_matrix;

typedef struct {
    slosh_matrix affine;
    float velocity[3];
    float pad0;
    float force_dt[3];
    float vel_grad_det;
    float mass;
    uint32_t enabled;
    uint32_t pad1[2];
} slosh_kinematics;

typedef struct {
    float init_volume;
    float init_radius;
    float damping;
    float phase;
    uint32_t fixed;
} slosh_particle_properties;

typedef struct { uint32_t words[13]; } slosh_particle_model;

/* The SoA buffers GpuParticles::from_particles uploads (src/solver/particle.rs:474-500),
 * i.e. SoAParticles::new applied to &[Particle<ParticleModel>] (particle.rs:401-429).
 * `scene_ids` (optional, batched mode): scene index per particle, < SLOSH_MAX_SCENES;
 * each scene then owns its own block-key space (its particles never share a block
 * with another scene's). */
typedef struct {
    size_t len;
    const float *positions; /* len x 4 floats (w ignored) */
    const slosh_kinematics *kinematics;
    const slosh_matrix *def_grad;
    const slosh_particle_properties *properties;
    const slosh_particle_model *models;
    const uint32_t *scene_ids;
} slosh_particles;

/* GpuBoundaryCondition, src/solver/boundary_condition.rs:6-10. */
typedef struct { uint32_t ty; float friction; } slosh_boundary_condition;
/* GpuShape, src/rbd/shapes.rs:66-70 (a.w = bit-cast shape tag). */
typedef struct { float a[4]; float b[4]; } slosh_shape;
/* GpuSim3 pose: rotation quaternion (x, y, z, w), translation xyz + uniform scale. */
typedef struct { float rotation[4]; float translation_scale[4]; } slosh_sim3;
/* GpuVelocity, src/rbd/dynamics/body.rs:24-32. */
typedef struct { float linear[3]; float pad0; float angular[3]; float pad1; } slosh_velocity;
/* GpuMassProperties, src/rbd/dynamics/body.rs:34-44. */
typedef struct {
    float inv_inertia[3][4];
    float inv_mass[3];
    float pad0;
    float com[3];
    float pad1;
} slosh_mass_properties;
/* BodyDesc, src/rbd/dynamics/body.rs:93-106. The caller performs rapier's
 * conversion (GpuBodySet::from_rapier, body.rs:183-235): fixed bodies and
 * one-way coupling get zero mass properties. */
typedef struct {
    slosh_mass_properties local_mprops;
    slosh_mass_properties mprops;
    slosh_velocity vel;
    slosh_sim3 pose;
    slosh_shape shape;
} slosh_body_desc;

/* One active block (GpuActiveBlockHeader, src/grid/grid.rs:335-346): virtual id,
 * first sorted slot, primary count, count including the "extras" of the +1
 * neighbourhood, and the 10 within-block slab-bucket counts (4 primary + 6 extra,
 * sort.slang:15-42). */
typedef struct {
    int32_t vid[3];
    uint32_t first_particle;
    uint32_t num_particles;
    uint32_t num_particles_with_extras;
    uint32_t bucket_counts[10];
} slosh_block_info;

/* MpmPipelineHooks, src/pipeline.rs:146-238. Every hook is optional (NULL = the
 * no-op `()` impl, pipeline.rs:240). Hooks enqueue on `stream` and return 0 or a
 * negative code. `run_p2g` / `run_g2p` set *handled = 1 to replace the built-in
 * pass. A run_p2g replacement writes node momentum/mass (slosh_device_view.nodes);
 * a run_g2p replacement writes velocity, velocity gradient and vel_grad_det. */
typedef int (*slosh_hook_fn)(slosh_data *data, slosh_stream stream, void *state);
typedef int (*slosh_replace_fn)(slosh_data *data, slosh_stream stream, void *state, int *handled);
typedef struct {
    slosh_hook_fn after_particle_sort;
    slosh_replace_fn run_p2g;
    slosh_replace_fn run_g2p;
    slosh_hook_fn after_p2g;
    slosh_hook_fn after_grid_update;
    slosh_hook_fn after_g2p;
    slosh_hook_fn after_particles_update;
} slosh_hooks;

/* Device-side view of the state, for hooks (layouts documented in slosh_api.hip). */
typedef struct {
    size_t num_particles;
    size_t particle_capacity;
    uint32_t grid_capacity;
    float cell_width;
    void *positions;     /* float4[cap]: xyz, w = enabled (u32 bits) -- current order   */
    void *vel_mass;      /* float4[cap]: velocity xyz, mass                              */
    void *affine;        /* float[9][cap] component-major shader matrix                 */
    void *def_grad;      /* float[9][cap]                                               */
    void *vel_grad_det;  /* float[cap]                                                  */
    void *particle_ids;  /* uint32[cap]: original particle index of each slot           */
    void *force_dt;      /* float[3][cap], indexed by current slot                      */
    void *nodes;         /* float4[grid_capacity*64] by header id: (velocity|momentum).xyz, mass */
    void *sorted_ids;    /* uint32[cap]: particle slots grouped by block and cell (substeps  */
                         /* that run hooks sort fully; others leave it stale)               */
    void *counters;      /* uint32[8]: [0] active blocks                                */
    void *block_list;    /* uint32[grid_capacity]: header ids of the active blocks (a     */
                         /* block's header id is its slot in the hash table), [0, counters[0]) */
} slosh_device_view;

/* ---- Pipeline: MpmPipeline::new / new_with_kernels (src/pipeline.rs:456-534). */
int slosh_pipeline_create(const slosh_kernels *kernels, slosh_pipeline **out);
void slosh_pipeline_destroy(slosh_pipeline *pipeline);
void slosh_kernels_default(slosh_kernels *out); /* MpmPipelineKernels::default, pipeline.rs:69-80 */

/* ---- State: MpmData::new / with_select_coupling (src/pipeline.rs:311-430).
 * Copies every input; the library owns the device memory. `grid_capacity` is
 * rounded up to a power of two (GpuGrid::with_capacity, src/grid/grid.rs:435-437).
 * `particle_capacity` (>= particles->len, 0 = len) reserves room for appends. */
int slosh_data_create(slosh_pipeline *pipeline, const slosh_sim_params *params,
                      const slosh_particles *particles, size_t particle_capacity,
                      const slosh_body_desc *bodies, size_t num_bodies,
                      const slosh_boundary_condition *body_materials, float cell_width,
                      uint32_t grid_capacity, slosh_data **out);
void slosh_data_destroy(slosh_data *data);

/* ---- Trimesh colliders: the body set's shared collision buffers
 * (GpuBodySet::shapes_collision_vertices / _indices, src/rbd/dynamics/body.rs:84-90),
 * as convert_trimesh_to_gpu (src/trimesh.rs:22-103) fills them: `vertices4` holds
 * one Vect per entry as x, y, z, pad (the BVH node AABBs, the mesh vertices, the
 * vertex and edge pseudo-normals), `indices` the BVH topology (entry, exit, shape
 * per node) then the triangle vertex indices. Bodies whose shape is a trimesh
 * (GpuShape::trimesh, src/rbd/shapes.rs:157-173) index into them. Copies both;
 * every trimesh body's ranges, triangle indices and BVH links are validated (links
 * must point forward, so the walk of collide() always ends): SLOSH_E_INVALID
 * otherwise. A data with trimesh bodies and no buffers fails at the first step. */
int slosh_data_set_collision_mesh(slosh_data *data, const float *vertices4, size_t num_vertices,
                                  const uint32_t *indices, size_t num_indices);

/* ---- One substep: MpmPipeline::launch_step (src/pipeline.rs:562-770).
 * Enqueues on `stream` and never synchronizes; `hooks` and `timestamps` may be NULL. */
int slosh_launch_step(slosh_pipeline *pipeline, slosh_data *data, slosh_stream stream,
                      const slosh_hooks *hooks, void *hooks_state, slosh_timestamps *timestamps);

/* Enqueues `count` substeps through a captured HIP graph (no hooks, no
 * timestamps): the substep command list is captured once per data and replayed. */
int slosh_launch_steps_graph(slosh_pipeline *pipeline, slosh_data *data, slosh_stream stream,
                             uint32_t count);

/* A captured sequence of substeps (no reference counterpart: wgpu command buffers
 * are recorded per frame; this is their hipGraph equivalent). `steps` must be even
 * (the substep alternates particle buffers and hash tables, so the state after an
 * even count matches the capture state); capture after at least one
 * slosh_launch_step. A graph is invalidated by any mutation of `data` (launch then
 * fails with SLOSH_E_INVALID). */
typedef struct slosh_graph slosh_graph;
int slosh_graph_capture(slosh_pipeline *pipeline, slosh_data *data, uint32_t steps, slosh_graph **out);
int slosh_graph_launch(slosh_graph *graph, slosh_stream stream);
void slosh_graph_destroy(slosh_graph *graph);

/* Precision of the SVD-based steps (fixed-corotated stress, Drucker-Prager
 * return mapping): 1 = FP64 (default), 0 = FP32 as the reference computes them.
 * No reference counterpart: the reference is FP32 only. */
int slosh_set_svd_fp64(slosh_data *data, int enable);

/* Rewrites SimulationParams (cf. src_testbed/step.rs:136-147). */
int slosh_set_sim_params(slosh_data *data, const slosh_sim_params *params);
/* WgTimestepBounds::compute_bounds (src/solver/timestep_bound.rs:56-79): the
 * largest stable timestep estimate over the enabled particles, in seconds
 * (computed_dt of the u32 fixed point in 1e-12 s units, also returned in raw if
 * non-null). Synchronous on `stream`. */
int slosh_timestep_bound(slosh_data *data, slosh_stream stream, float *out_dt, uint32_t *out_raw);
/* No counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops). slosh_timestep_bound's u32 fixed-point
 * bound, bit-identical, written to DEVICE memory `out_raw` (0xffffffff with no
 * enabled particle), asynchronous on `stream`: no allocation, no memset, no copy,
 * no synchronization. Its partial minima (one per 2048 particles) live in a
 * static 64 KiB array per GPU, so at most SLOSH_NOALLOC_MAX_PARTICLES particles
 * (else SLOSH_E_INVALID). Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams, no ordering between them) race. */
#define SLOSH_NOALLOC_MAX_PARTICLES 33554432u
int slosh_timestep_bound_noalloc(slosh_data *data, slosh_stream stream, uint32_t *out_raw);
/* The adaptive substep count of src_testbed/step.rs:100-147: with
 * min_substeps < max_substeps, ceil(base_dt / bound) clamped to the range;
 * otherwise max_substeps. Sets the simulation dt to base_dt / count. */
int slosh_adaptive_substeps(slosh_data *data, slosh_stream stream, float base_dt, uint32_t min_substeps,
                            uint32_t max_substeps, uint32_t *out_substeps);

/* Synchronizes `stream`, returns SLOSH_E_KEY_RANGE if any substep since the last
 * call saw a particle outside the block-key range, else SLOSH_E_GRID_OVERFLOW if
 * any overflowed the sparse-grid hash map, and clears the flags. */
int slosh_check_errors(slosh_data *data, slosh_stream stream);

/* ---- Readback (staging copies of src_testbed/step.rs:271-334), in the
 * original particle order, reference layouts. Synchronous on `stream`. */
size_t slosh_num_particles(const slosh_data *data);
int slosh_read_positions(slosh_data *data, slosh_stream stream, float *out4);
int slosh_read_kinematics(slosh_data *data, slosh_stream stream, slosh_kinematics *out);
int slosh_read_def_grad(slosh_data *data, slosh_stream stream, slosh_matrix *out);
int slosh_read_models(slosh_data *data, slosh_stream stream, slosh_particle_model *out);
int slosh_read_properties(slosh_data *data, slosh_stream stream, slosh_particle_properties *out);
int slosh_read_scene_ids(slosh_data *data, slosh_stream stream, uint32_t *out);
/* Body poses/velocities after the last integrate_bodies pass (poses_staging, pipeline.rs:270). */
int slosh_read_bodies(slosh_data *data, slosh_stream stream, slosh_sim3 *poses, slosh_velocity *vels);
/* Active blocks of the last substep's sort (any order; compare as sets keyed by vid). */
int slosh_num_blocks(slosh_data *data, slosh_stream stream, uint32_t *out);
int slosh_read_blocks(slosh_data *data, slosh_stream stream, slosh_block_info *out, uint32_t max_blocks,
                      uint32_t *num_written);
/* Grid nodes after the last grid update, keyed like slosh_read_blocks: out[b*64+n] = (v.xyz, m). */
int slosh_read_nodes(slosh_data *data, slosh_stream stream, float *out4, uint32_t max_blocks);
/* Scene ids of the blocks returned by slosh_read_blocks (batched mode; 0 otherwise). */
int slosh_read_block_scenes(slosh_data *data, slosh_stream stream, uint32_t *out, uint32_t max_blocks);

/* ---- Mutation: GpuParticles::append / shift_remove (src/solver/particle.rs:505-572).
 * Appending within the capacity given at creation uploads only the new particles;
 * beyond it the storage is reallocated. Removal runs on the device (order-preserving
 * compaction; the remaining particles keep their relative creation order and are
 * renumbered from 0). Both are enqueued on `stream` and invalidate captured graphs.
 * An append within the capacity copies `particles` into a pinned host arena owned
 * by `data` before it returns (the caller's arrays may be reused at once) and
 * leaves the uploads queued on `stream`, ahead of the substeps launched next. */
int slosh_append_particles(slosh_data *data, slosh_stream stream, const slosh_particles *particles);
/* No reference counterpart (serving): pins `bytes` of host memory as a ring that
 * later appends stage in (200 bytes per particle), so that admitting a scene pins
 * nothing; an append that does not fit the ring falls back to the arenas. An
 * append takes the region after the previous one and waits only while the ring is
 * full (for the oldest region's uploads). 0 releases the ring. Synchronizes with
 * the uploads still reading the previous ring. */
int slosh_reserve_staging(slosh_data *data, size_t bytes);
int slosh_remove_particles(slosh_data *data, slosh_stream stream, size_t start, size_t end);
/* Removes every particle of `scene_id` / of the listed scenes (batched mode; no
 * reference counterpart: scenes are this library's batching unit). */
int slosh_remove_scene(slosh_data *data, slosh_stream stream, uint32_t scene_id, size_t *removed);
int slosh_remove_scenes(slosh_data *data, slosh_stream stream, const uint32_t *scene_ids, size_t count,
                        size_t *removed);
/* Number of particles of a scene, and their state in creation order (positions as
 * float4 with w = 0; kinematics / def_grad in reference layouts; any output may be
 * null). Gathered on the device; synchronous on `stream`. */
size_t slosh_scene_size(const slosh_data *data, uint32_t scene_id);
int slosh_read_scene(slosh_data *data, slosh_stream stream, uint32_t scene_id, float *positions4,
                     slosh_kinematics *kinematics, slosh_matrix *def_grad, size_t max_count, size_t *count);
/* Overwrites kinematics (e.g. force_dt) in the original particle order. */
int slosh_write_kinematics(slosh_data *data, slosh_stream stream, const slosh_kinematics *kin);

int slosh_device_view_get(slosh_data *data, slosh_device_view *out);

/* ---- Launch shape (no reference counterpart: the wgpu pipeline fixes its
 * workgroup sizes in the shaders). p2g_chunk: particles staged in LDS per P2G
 * round (192, 256, 384, 512); g2p_threads: threads per G2P/particle-update
 * workgroup (128, 256); block_wgs: workgroups of the per-block kernels, which
 * loop over the active blocks (64..65536). Results do not depend on it beyond
 * summation order within a cell. slosh_autotune times `trial_steps` substeps of
 * a private copy of `data` per candidate (coordinate search, best of two
 * replays), keeps the fastest in `data` and returns it in `out` (may be null);
 * `data` itself does not advance. */
typedef struct slosh_tuning {
    uint32_t p2g_chunk;
    uint32_t g2p_threads;
    uint32_t block_wgs;
} slosh_tuning;
int slosh_set_tuning(slosh_data *data, const slosh_tuning *tuning);
int slosh_get_tuning(const slosh_data *data, slosh_tuning *out);
int slosh_autotune(slosh_pipeline *pipeline, slosh_data *data, slosh_stream stream, uint32_t trial_steps,
                   slosh_tuning *out);

/* ---- Timing: per-pass GPU timestamps (GpuTimestamps; pass labels of
 * src/pipeline.rs:588-757: grid_sort, p2g, grid_update, g2p, particles_update,
 * integrate_bodies). The timestamps are written by the dispatch packets of each
 * pass's first and last kernel. The fused G2P + particle-update kernel (which also
 * integrates the rigid bodies) is reported as g2p. */
#define SLOSH_NUM_PASSES 6
int slosh_timestamps_create(slosh_timestamps **out);
void slosh_timestamps_destroy(slosh_timestamps *ts);
/* Milliseconds per pass accumulated since the last reset (synchronizes). */
int slosh_timestamps_read(slosh_timestamps *ts, double *ms_per_pass, uint32_t *count);
void slosh_timestamps_reset(slosh_timestamps *ts);

/* ---- Streams and events (plumbing for callers without a HIP binding). */
int slosh_stream_create(slosh_stream *out);
int slosh_stream_destroy(slosh_stream stream);
int slosh_stream_synchronize(slosh_stream stream);
int slosh_event_create(void **out);
int slosh_event_destroy(void *event);
int slosh_event_record(void *event, slosh_stream stream);
int slosh_event_elapsed_ms(void *start, void *stop, float *ms);
int slosh_device_count(int *out);

const char *slosh_last_error(void);
const char *slosh_version(void);

#ifdef __cplusplus
}
#endif
#endif
