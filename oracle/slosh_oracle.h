/*
 * slosh_oracle.h -- CPU oracle for the slosh MPM substep (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C, single-precision restatement of the reference substep
 * `MpmPipeline::launch_step` (src/pipeline.rs:562-770) and of every compute
 * shader it dispatches, written from the shader sources read as text. It is
 * the checker for the HIP implementation in slosh_amd/: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the measured or shipped product.
 *
 * Atomic operations of the reference (hash-map CAS, block counters, bucket
 * cursors) are executed sequentially in thread-index order, so the oracle's
 * header ids and within-bucket particle orders are one of the orders the
 * reference can produce. Byte layouts follow the reference's host structs:
 *   Kinematics 96 B (src/solver/particle.rs:132-147; shaders/slosh/solver/particle.slang:12-27)
 *   ParticleProperties 20 B (src/solver/particle.rs:154-167)
 *   GpuParticleModel 52 B (src/solver/particle_model.rs:87-102)
 *   Matrix = 3 x vec4, row r of the shader matrix == glam column r.
 */
#ifndef SLOSH_ORACLE_H
#define SLOSH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float m[3][4]; } so_mat3;

typedef struct {
    so_mat3 affine;        /* @0  */
    float velocity[3];     /* @48 */
    float pad0;
    float force_dt[3];     /* @64 */
    float vel_grad_det;    /* @76 */
    float mass;            /* @80 */
    uint32_t enabled;      /* @84 */
    uint32_t pad1[2];      /* 96 B */
} so_kinematics;

typedef struct {
    float init_volume, init_radius, damping, phase;
    uint32_t fixed;
} so_props;

typedef struct { uint32_t w[13]; } so_model; /* tag @0, payload @4 (52 B) */

typedef struct { float gravity[3]; float dt; } so_sim_params;
typedef struct { uint32_t ty; float friction; } so_bc;
typedef struct { float a[4]; float b[4]; } so_shape;
/* Trimesh collision buffers (convert_trimesh_to_gpu, src/trimesh.rs:22-103):
 * vtx = x, y, z, pad per entry; idx = BVH topology then triangle indices. */
typedef struct { const float *vtx; const uint32_t *idx; size_t nv, ni; } so_mesh;
typedef struct { float rotation[4]; float translation_scale[4]; } so_sim3; /* quat xyzw; t.xyz, scale */
typedef struct { float linear[3]; float pad0; float angular[3]; float pad1; } so_velocity;
typedef struct { float inv_inertia[3][4]; float inv_mass[3]; float pad0; float com[3]; float pad1; } so_mprops;
typedef struct {
    so_mprops local_mprops;
    so_mprops mprops;
    so_velocity vel;
    so_sim3 pose;
    so_shape shape;
} so_body;

typedef struct {
    int32_t vid[3];
    uint32_t first_particle;
    uint32_t num_particles;
    uint32_t num_particles_with_extras;
    uint32_t bucket_counts[10];
} so_block_info;

typedef struct so_state so_state;

/* svd_mode: 0 = every SVD-based step in FP32 (as the reference computes it);
 * 1 = the SVD, the fixed-corotated F - R assembly and the Drucker-Prager return
 * mapping in FP64, rounded once to FP32. */
so_state *so_create(const so_sim_params *params, size_t n, const float *pos4,
                    const so_kinematics *kin, const so_mat3 *def_grad,
                    const so_props *props, const so_model *models, size_t nbodies,
                    const so_body *bodies, const so_bc *materials, float cell_width,
                    uint32_t grid_capacity, int svd_mode);
void so_destroy(so_state *s);
/* One substep. Returns 0, or -1 on hash-map overflow. */
int so_step(so_state *s);
void so_set_threads(int nthreads);
/* P2G formulation (process-wide): 0 = block-local scatter (default), 1 = the
 * shader's per-node gather (p2g_scatter_style.slang:41-155). Bit-identical. */
void so_set_p2g_gather(int gather);

void so_read_positions(const so_state *s, float *out4);
void so_read_kinematics(const so_state *s, so_kinematics *out);
void so_read_def_grad(const so_state *s, so_mat3 *out);
void so_read_models(const so_state *s, so_model *out);
void so_read_bodies(const so_state *s, so_sim3 *poses, so_velocity *vels);
uint32_t so_num_blocks(const so_state *s);
void so_read_blocks(const so_state *s, so_block_info *out);
/* Grid nodes after the grid update of the last step: out[b*64+n] = (v.xyz, m). */
void so_read_nodes(const so_state *s, float *out4);

/* Stand-alone pieces, exposed for known-answer tests. */
void so_svd3(const double a[9], int f64, double u[9], double sig[3], double v[9]);
/* Applies the material update of default.slang:51-92 to def_grad (row-major 3x3,
 * shader convention) and returns the Kirchhoff stress; model is updated in place. */
void so_model_update(so_model *model, float def_grad[9], float stress[9], int svd_mode);
/* Timestep bound (WgTimestepBounds): min over enabled particles of the sound-speed
 * and velocity bounds, as the reference's u32 fixed point in 1e-12 s units. */
uint32_t so_timestep_bound(const so_state *s);
uint32_t so_particle_timestep_bound(const so_model *model, const so_kinematics *kin, const so_mat3 *def_grad,
                                    const so_props *props, float cell_width);
void so_prefix_sum(uint32_t *v, size_t len);
/* Grid update of one node at cell_pos (update_single_cell + collide + project_velocity),
 * mvm = (momentum.xyz, mass) in, (velocity.xyz, mass) out; out_normal_dist = (n, dist). */
void so_grid_update_node(const so_body *bodies, const so_bc *mats, size_t nb, float h, const so_sim_params *params,
                         const float cell_pos[3], float mvm[4], float out_normal_dist[4]);
uint32_t so_pack_key(int32_t x, int32_t y, int32_t z);
/* Shape::project_point_on_boundary (shape.slang:180-201) for one collider shape
 * and pose; returns 0 for the shapes collide() skips (polyline, trimesh). */
int so_project_point_on_boundary(const so_shape *sh, const so_sim3 *pose, const float pt[3], const so_mesh *mesh,
                                 float max_dist, float out[3], int *inside);
/* Shared trimesh collision buffers of the body set (copied). */
void so_set_collision_mesh(so_state *s, const float *vtx4, size_t num_vertices, const uint32_t *idx,
                           size_t num_indices);
uint32_t so_hash(uint32_t packed_key);

#ifdef __cplusplus
}
#endif
#endif
