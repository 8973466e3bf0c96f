"""Seeded scene generators for the benchmark configurations and the parity tests.

RNG: SplitMix64 with ``u = (x >> 40) * 2**-24``; the seed of configuration k is
``0x5105_0000 + k`` (batched scenes add the scene index). The generators are
shared by the HIP path and by the CPU checker so both see identical inputs.

Configurations (sizes as in the benchmark definition):

* ``elastic_cube``   -- 32^3 lattice, spacing 1/128, cube of side 0.25 centred at
  (0.5, 0.4, 0.5), jitter U(+-0.1 dx) with dx = 1/64, neo-Hookean E=1e5 nu=0.3,
  rho=1000, r = dx/4, six sticky cuboid walls, dt=1e-4.
* ``dam_break``      -- 2^20 particles on a 64 x 256 x 64 lattice at spacing 1/512
  (column [0,0.125] x [0,0.5] x [0,0.125]), dx = 1/256, neo-Hookean with
  lambda=1e5, mu=0 (weakly compressible, bulk modulus 1e5), rho=1000,
  separating walls with zero friction, dt=2e-5.
* ``sand_spheres``   -- 2^22 particles, dx=1/512: a Drucker-Prager sand bed
  (friction angle 30 deg, E=3.5e5, nu=0.3, rho=1600) plus 64 neo-Hookean spheres
  (E=2e5, nu=0.4, R=0.03, rho=1000) with velocities U[-0.5,0.5]^3 and eight
  two-way rigid boxes of side 0.1 at seeded poses, dt=2e-5.
* ``batched_scene``  -- one scene of the 256-scene batch: N log-uniform in
  [2e3, 2e5], material uniform over NH / FC / DP, E log-uniform [1e4, 1e6],
  nu U[0.2, 0.45], dx = 1/128, the elastic-cube layout scaled to N.
* ``served_request`` -- one request of the served-scenes stream: a batched-style
  scene with N log-uniform in [5e3, 2e6] and substeps uniform in [100, 2000];
  ``poisson_arrivals`` gives the arrival times.
"""
import numpy as np

from . import layouts as L
from . import models as M

SEED_BASE = 0x51050000
_MASK = (1 << 64) - 1


def splitmix64(seed, count):
    """``count`` consecutive SplitMix64 outputs for ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _MASK) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, count):
    """U[0,1) doubles from the top 24 bits of each SplitMix64 output."""
    return (splitmix64(seed, count) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


class Scene:
    """Inputs of ``MpmData::new``: params, particles (reference SoA layouts),
    bodies + per-body boundary conditions, cell width and grid capacity."""

    def __init__(self, name, particles, dt, cell_width, grid_capacity, bodies, materials,
                 gravity=(0.0, -9.81, 0.0), substeps=1):
        self.name = name
        self.particles = particles
        self.params = np.zeros(1, dtype=L.SIM_PARAMS)[0]
        self.params["gravity"] = gravity
        self.params["dt"] = dt
        self.cell_width = np.float32(cell_width)
        self.grid_capacity = int(grid_capacity)
        self.bodies = np.asarray(bodies, dtype=L.BODY).reshape(-1)
        self.materials = np.asarray(materials, dtype=L.BOUNDARY_COND).reshape(-1)
        self.substeps = substeps

    @property
    def n(self):
        return self.particles["pos4"].shape[0]


def box_walls(bc_type, friction=0.0, thickness=0.1):
    """Six fixed cuboid colliders enclosing [0,1]^3 (fixed bodies: zero mass props)."""
    t = thickness / 2
    ext = 0.5 + thickness
    specs = [((0.5, -t, 0.5), (ext, t, ext)), ((0.5, 1.0 + t, 0.5), (ext, t, ext)),
             ((-t, 0.5, 0.5), (t, ext, ext)), ((1.0 + t, 0.5, 0.5), (t, ext, ext)),
             ((0.5, 0.5, -t), (ext, ext, t)), ((0.5, 0.5, 1.0 + t), (ext, ext, t))]
    bodies = [M.body(M.cuboid(he), c) for c, he in specs]
    mats = [M.boundary_condition(bc_type, friction)] * 6
    return bodies, mats


def lattice(counts, spacing, origin, jitter, seed):
    """Regular lattice (x fastest, then y, then z) with seeded uniform jitter."""
    nx, ny, nz = counts
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    zz, yy, xx = np.meshgrid(k, j, i, indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    pts = (pts + 0.5) * spacing + np.asarray(origin, dtype=np.float64)
    u = uniform(seed, pts.shape[0] * 3).reshape(-1, 3)
    pts += (2.0 * u - 1.0) * jitter
    return pts.astype(np.float32)


def elastic_cube(n_side=32, seed=SEED_BASE + 0):
    dx = 1.0 / 64
    spacing = 1.0 / 128
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.4 - side / 2, 0.5 - side / 2)
    pos = lattice((n_side,) * 3, spacing, origin, 0.1 * dx, seed)
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, M.elastic_neo_hookean(1.0e5, 0.3))
    bodies, mats = box_walls(M.STICK)
    return Scene("elastic_cube", parts, 1.0e-4, dx, 1024, bodies, mats, substeps=100)


def dam_break(counts=(64, 256, 64), seed=SEED_BASE + 1, grid_capacity=16384):
    dx = 1.0 / 256
    spacing = 1.0 / 512
    pos = lattice(counts, spacing, (0.0, 0.0, 0.0), 0.1 * dx, seed)
    model = M.elastic_from_lame(M.ELASTIC_NEO_HOOKEAN, 1.0e5, 0.0)
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, model)
    bodies, mats = box_walls(M.SEPARATE, 0.0)
    return Scene("dam_break", parts, 2.0e-5, dx, grid_capacity, bodies, mats, substeps=1000)


def _ball_points(n, radius, center, seed):
    u = uniform(seed, n * 3).reshape(-1, 3)
    r = radius * np.cbrt(u[:, 0])
    ct = 2.0 * u[:, 1] - 1.0
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2.0 * np.pi * u[:, 2]
    d = np.stack([st * np.cos(ph), ct, st * np.sin(ph)], axis=1)
    return (np.asarray(center) + d * r[:, None]).astype(np.float32)


def sand_spheres(n_sand=2097152, n_per_sphere=32768, seed=SEED_BASE + 2, grid_capacity=1 << 17, n_spheres=64):
    dx = 1.0 / 512
    u = uniform(seed, n_sand * 6).reshape(-1, 6)
    lo = np.array([0.375, 0.0, 0.375])
    ext = np.array([0.25, 0.03125, 0.25])
    sand_pos = (lo + u[:, :3] * ext).astype(np.float32)
    sand_vel = (u[:, 3:] - 0.5).astype(np.float32)
    sand = M.particle_arrays(sand_pos, np.float32(dx / 4), 1600.0, M.sand(3.5e5, 0.3, 30.0), sand_vel)
    parts = [sand]
    nh = M.elastic_neo_hookean(2.0e5, 0.4)
    for s in range(n_spheres):
        i, j, k = s % 4, (s // 4) % 4, s // 16
        c = (0.5 + (i - 1.5) * 0.07, 0.09 + j * 0.07, 0.5 + (k - 1.5) * 0.07)
        sp = _ball_points(n_per_sphere, 0.03, c, seed * 131 + s + 1)
        sv = (uniform(seed * 257 + s + 1, n_per_sphere * 3).reshape(-1, 3) - 0.5).astype(np.float32)
        parts.append(M.particle_arrays(sp, np.float32(dx / 4), 1000.0, nh, sv))
    particles = M.concat_particles(parts)
    bodies, mats = box_walls(M.SEPARATE, 0.0)
    ub = uniform(seed * 7 + 3, 8 * 6).reshape(8, 6)
    for b in range(8):
        t = (0.2 + 0.6 * ub[b, 0], 0.4 + 0.2 * ub[b, 1], 0.2 + 0.6 * ub[b, 2])
        rot = M.quat_from_scaled_axis((ub[b, 3:] - 0.5) * 2.0)
        bodies.append(M.body(M.cuboid((0.05, 0.05, 0.05)), t, rot, mass=1.0, inv_inertia_diag=(600.0,) * 3))
        mats.append(M.boundary_condition(M.SEPARATE, 0.5))
    return Scene("sand_spheres", particles, 2.0e-5, dx, grid_capacity, bodies, mats, substeps=500)


def batched_scene(scene_index, seed_base=SEED_BASE + 3, n_override=None):
    """One scene of the 256-scene batch (elastic-cube layout scaled to N)."""
    seed = seed_base + scene_index
    u = uniform(seed * 1009 + 17, 4)
    n = int(np.exp(np.log(2.0e3) + u[0] * (np.log(2.0e5) - np.log(2.0e3)))) if n_override is None else n_override
    mat = int(u[1] * 3.0) % 3
    e = float(np.exp(np.log(1.0e4) + u[2] * (np.log(1.0e6) - np.log(1.0e4))))
    nu = 0.2 + 0.25 * u[3]
    dx = 1.0 / 128
    side = int(np.ceil(n ** (1.0 / 3.0)))
    spacing = dx / 2
    ext = side * spacing
    origin = (0.5 - ext / 2, 0.4 - ext / 2, 0.5 - ext / 2)
    pos = lattice((side, side, side), spacing, origin, 0.1 * dx, seed)[:n]
    model = [M.elastic_neo_hookean(e, nu), M.elastic(e, nu), M.sand(e, nu, 30.0)][mat]
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, model)
    bodies, mats = box_walls(M.STICK)
    return Scene(f"batched_{scene_index}", parts, 1.0e-4, dx, 4096, bodies, mats, substeps=200)


def served_request(k, seed_base=SEED_BASE + 4, n_range=(5.0e3, 2.0e6), substep_range=(100, 2000)):
    """Request k of the served-scenes stream: a batched-configuration scene with N
    log-uniform in n_range and a substep count uniform in substep_range.
    Returns (scene, substeps)."""
    u = uniform(seed_base * 7919 + k, 2)
    n = int(np.exp(np.log(n_range[0]) + u[0] * (np.log(n_range[1]) - np.log(n_range[0]))))
    steps = int(substep_range[0] + u[1] * (substep_range[1] - substep_range[0] + 1))
    steps = min(steps, substep_range[1])
    sc = batched_scene(k, seed_base=seed_base, n_override=n)
    sc.name = f"request_{k}"
    return sc, steps


def poisson_arrivals(rate, count, seed=SEED_BASE + 4):
    """Arrival times (s) of a Poisson process of the given rate (exponential gaps)."""
    u = uniform(seed * 104729 + int(rate * 1000), count)
    return np.cumsum(-np.log1p(-u) / rate)


def stir(scene, seed, rate=2.0):
    """Gives a scene a smooth nonuniform initial velocity field, v = w x r + A r
    about its centroid with w and A uniform in [-rate, rate] (1/s), so that the
    APIC matrices and deformation gradients it develops are well above round-off
    (a scene falling from rest has a uniform velocity field and C ~ 0).
    Returns the scene, modified in place."""
    pos = scene.particles["pos4"][:, :3].astype(np.float64)
    r = pos - pos.mean(axis=0)
    u = uniform(seed, 12)
    w = (2.0 * u[:3] - 1.0) * rate
    a = ((2.0 * u[3:] - 1.0) * rate).reshape(3, 3)
    v = np.cross(w, r) + r @ a.T
    scene.particles["kin"]["velocity"] = v.astype(np.float32)
    return scene


def estimate_blocks(pos, cell_width):
    """Upper estimate of the active 4x4x4 blocks a particle cloud touches: its
    bounding box in blocks plus the +1 neighbour shell on every axis."""
    if len(pos) == 0:
        return 0
    lo = np.floor(np.min(pos[:, :3], axis=0) / (4.0 * cell_width)) - 1
    hi = np.floor(np.max(pos[:, :3], axis=0) / (4.0 * cell_width)) + 2
    return int(np.prod(hi - lo + 1))


def pow2_at_least(v):
    c = 1
    while c < v:
        c <<= 1
    return c


def _values(rec):
    """Field values of a structured record or array (padding bytes excluded)."""
    return repr(np.asarray(rec).reshape(-1).tolist())


class Batch:
    """Independent scenes stepped together: every particle of scenes[k] carries
    scene id ids[k] (its own block-key space); time step, gravity, cell width and
    colliders are shared, as one launch applies them to every scene."""

    def __init__(self, scene_list, ids=None, grid_capacity=None):
        assert scene_list, "empty batch"
        s0 = scene_list[0]
        for sc in scene_list[1:]:
            assert _values(sc.params) == _values(s0.params), "scenes of a batch share time step and gravity"
            assert sc.cell_width == s0.cell_width, "scenes of a batch share the cell width"
            assert _values(sc.bodies) == _values(s0.bodies) and _values(sc.materials) == _values(s0.materials), \
                "scenes of a batch share their colliders"
        self.scenes = list(scene_list)
        self.ids = np.arange(len(scene_list), dtype=np.uint32) if ids is None else np.asarray(ids, np.uint32)
        self.counts = np.array([sc.n for sc in scene_list], dtype=np.int64)
        self.particles = M.concat_particles([sc.particles for sc in scene_list])
        self.scene_ids = np.repeat(self.ids, self.counts).astype(np.uint32)
        self.params = s0.params
        self.cell_width = s0.cell_width
        self.bodies = s0.bodies
        self.materials = s0.materials
        if grid_capacity is None:
            est = sum(estimate_blocks(sc.particles["pos4"], sc.cell_width) for sc in scene_list)
            grid_capacity = pow2_at_least(est + est // 4 + 1024)
        self.grid_capacity = int(grid_capacity)
        self.name = f"batch of {len(scene_list)}"

    @property
    def n(self):
        return int(self.counts.sum())


def shapes_test_scene(n_side=18, seed=4321, moving=True):
    """Parity scene for the capsule, cone and cylinder colliders (collide.slang with
    capsule.slang, cone.slang, cylinder.slang): a stirred neo-Hookean / fixed-corotated
    block thrown down onto a rotated capsule, a cone and a tilted, scaled cylinder inside
    the six walls; optionally the capsule moves (two-way body), which exercises the
    collision cache with these shapes."""
    dx = 1.0 / 64
    spacing = dx / 2
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.05, 0.5 - side / 2)  # the block encloses the three shapes
    pos = lattice((n_side,) * 3, spacing, origin, 0.1 * dx, seed)
    n = pos.shape[0]
    u = uniform(seed + 1, n * 3).reshape(n, 3)
    rel = pos.astype(np.float64) - np.array([0.5, 0.05 + side / 2, 0.5])
    vel = np.cross(np.array([0.6, -0.4, 0.9]), rel) + np.array([0.0, -0.6, 0.0]) + (u - 0.5) * 0.05
    tags = np.arange(n) % 2
    models = np.stack([M.elastic_neo_hookean(3.0e4, 0.3), M.elastic(2.5e4, 0.3)])[tags]
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, models, vel.astype(np.float32))
    bodies, mats = box_walls(M.SLIP)
    bodies.append(M.body(M.capsule((-0.05, 0.0, -0.02), (0.06, 0.01, 0.03), 0.02), (0.47, 0.1, 0.5),
                         M.quat_from_scaled_axis((0.4, -0.2, 0.7)),
                         linvel=(0.1, 0.0, -0.05) if moving else (0, 0, 0), angvel=(0.0, 1.0, 0.5) if moving else (0, 0, 0),
                         mass=0.4 if moving else None, inv_inertia_diag=(800.0, 800.0, 800.0) if moving else None))
    mats.append(M.boundary_condition(M.SEPARATE, 0.2))
    bodies.append(M.body(M.cone(0.04, 0.035), (0.55, 0.06, 0.46), M.quat_from_scaled_axis((0.0, 0.3, 0.0))))
    mats.append(M.boundary_condition(M.STICK))
    cyl = M.body(M.cylinder(0.03, 0.025), (0.52, 0.09, 0.56), M.quat_from_scaled_axis((1.0, 0.0, 0.4)))
    cyl["pose"]["translation_scale"][3] = 1.25
    bodies.append(cyl)
    mats.append(M.boundary_condition(M.SLIP))
    return Scene("shapes", parts, 1.0e-4, dx, 2048, bodies, mats, substeps=100)


def trimesh_test_scene(n_side=18, seed=2468, moving=True):
    """Parity scene for triangle-mesh colliders (collide.slang:120-130 with
    collision/trimesh.slang): a stirred neo-Hookean block around a rotated box mesh
    (moving two-way body when `moving`: collision cache) and a scaled icosphere mesh,
    inside the six walls. The mesh buffers ride on the scene as collision_vertices /
    collision_indices."""
    from . import trimesh as TM
    dx = 1.0 / 64
    spacing = dx / 2
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.05, 0.5 - side / 2)
    pos = lattice((n_side,) * 3, spacing, origin, 0.1 * dx, seed)
    n = pos.shape[0]
    u = uniform(seed + 1, n * 3).reshape(n, 3)
    rel = pos.astype(np.float64) - np.array([0.5, 0.05 + side / 2, 0.5])
    vel = np.cross(np.array([-0.5, 0.7, 0.3]), rel) + np.array([0.0, -0.5, 0.0]) + (u - 0.5) * 0.05
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, M.elastic_neo_hookean(3.0e4, 0.3),
                              vel.astype(np.float32))
    bodies, mats = box_walls(M.SLIP)
    buf = TM.CollisionBuffers()
    box = TM.convert(*TM.box_mesh((0.035, 0.02, 0.03)), buf)
    bodies.append(M.body(box, (0.47, 0.1, 0.49), M.quat_from_scaled_axis((0.5, 0.3, -0.4)),
                         linvel=(0.08, 0.0, 0.05) if moving else (0, 0, 0), angvel=(0.5, 0.0, 1.0) if moving else (0, 0, 0),
                         mass=0.4 if moving else None, inv_inertia_diag=(800.0, 800.0, 800.0) if moving else None))
    mats.append(M.boundary_condition(M.SEPARATE, 0.2))
    sph = TM.convert(*TM.icosphere(0.025, 1), buf)
    body = M.body(sph, (0.54, 0.08, 0.55), M.quat_from_scaled_axis((0.0, 0.4, 0.2)))
    body["pose"]["translation_scale"][3] = 1.2
    bodies.append(body)
    mats.append(M.boundary_condition(M.STICK))
    sc = Scene("trimesh", parts, 1.0e-4, dx, 2048, bodies, mats, substeps=100)
    sc.collision_vertices = buf.vertex_array()
    sc.collision_indices = buf.index_array()
    return sc


def heightfield_test_scene(n_side=18, seed=1357, sand=False):
    """Parity scene for heightfield colliders (3D heightfields collide as the
    trimesh of their samples, src/rbd/shapes.rs:252-268; cf. the reference's
    heightfield3 example): a block falling onto a rotated, bumpy heightfield with a
    stick boundary, inside the six walls. `sand`: Drucker-Prager sand instead of
    neo-Hookean."""
    from . import trimesh as TM
    dx = 1.0 / 64
    spacing = dx / 2
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.075, 0.5 - side / 2)
    pos = lattice((n_side,) * 3, spacing, origin, 0.1 * dx, seed)
    n = pos.shape[0]
    u = uniform(seed + 1, n * 3).reshape(n, 3)
    vel = np.array([0.1, -1.0, 0.05]) + (u - 0.5) * 0.05
    model = M.sand(1.0e5, 0.2) if sand else M.elastic_neo_hookean(3.0e4, 0.3)
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1500.0 if sand else 1000.0, model, vel.astype(np.float32))
    bodies, mats = box_walls(M.SLIP)
    i, j = np.meshgrid(np.arange(24), np.arange(20), indexing="ij")
    buf = TM.CollisionBuffers()
    hf = TM.convert(*TM.heightfield(np.sin(i / 3.0) * np.cos(j / 2.5), (0.4, 0.03, 0.45)), buf)
    bodies.append(M.body(hf, (0.5, 0.07, 0.5), M.quat_from_scaled_axis((0.05, 0.3, -0.04))))
    mats.append(M.boundary_condition(M.STICK))
    sc = Scene("heightfield_sand" if sand else "heightfield", parts, 1.0e-4, dx, 2048, bodies, mats, substeps=100)
    sc.collision_vertices = buf.vertex_array()
    sc.collision_indices = buf.index_array()
    return sc


def mixed_test_scene(n_side=20, seed=12345, moving_body=True):
    """Parity scene that exercises every branch: all four material tags, nonzero
    velocity gradients and affine matrices, perturbed deformation gradients, a
    fixed particle, damping, a rotated cuboid, a ball and (optionally) a moving
    two-way body so that the collision cache is exercised."""
    dx = 1.0 / 64
    spacing = dx / 2
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.12, 0.5 - side / 2)
    pos = lattice((n_side,) * 3, spacing, origin, 0.1 * dx, seed)
    n = pos.shape[0]
    u = uniform(seed + 1, n * 16).reshape(n, 16)
    rel = pos.astype(np.float64) - np.array([0.5, 0.12 + side / 2, 0.5])
    omega = np.array([0.3, 1.1, -0.4])
    vel = np.cross(omega, rel) + np.array([0.05, -0.2, 0.1]) + (u[:, :3] - 0.5) * 0.05
    tags = (np.arange(n) * 7 // 3) % 4
    recs = np.stack([M.elastic(2.0e4, 0.3), M.elastic_neo_hookean(3.0e4, 0.25),
                     M.sand(5.0e4, 0.3, 30.0), M.sand(4.0e4, 0.35, 25.0, neo_hookean=True)])
    models = recs[tags]
    parts = M.particle_arrays(pos, np.float32(dx / 4), 1000.0, models, vel.astype(np.float32))
    aff = (u[:, 3:12].reshape(n, 3, 3) - 0.5) * 0.02
    parts["kin"]["affine"] = L.mat_to_slang(aff)
    f = np.eye(3)[None] + (u[:, 7:16].reshape(n, 3, 3) - 0.5) * 0.01
    parts["def_grad"] = L.mat_to_slang(f)
    parts["props"]["damping"][::5] = 0.5
    parts["props"]["fixed"][::97] = 1
    parts["kin"]["force_dt"][::3] = (0.0, 1.0e-7, 0.0)
    bodies, mats = box_walls(M.SLIP)
    mats[0] = M.boundary_condition(M.SEPARATE, 0.3)
    mats[1] = M.boundary_condition(M.FRICTION_Z_UP, 0.2)
    mats[2] = M.boundary_condition(M.STICK)
    rot = M.quat_from_scaled_axis((0.2, 0.5, -0.3))
    bodies.append(M.body(M.cuboid((0.04, 0.03, 0.05)), (0.5, 0.08, 0.5), rot))
    mats.append(M.boundary_condition(M.SEPARATE, 0.1))
    bodies.append(M.body(M.ball(0.03), (0.44, 0.2, 0.47)))
    mats.append(M.boundary_condition(M.SLIP))
    if moving_body:
        bodies.append(M.body(M.cuboid((0.03, 0.03, 0.03)), (0.56, 0.2, 0.52), linvel=(-0.4, 0.1, 0.0),
                             angvel=(0.0, 2.0, 1.0), mass=0.5, inv_inertia_diag=(900.0, 900.0, 900.0)))
        mats.append(M.boundary_condition(M.SEPARATE, 0.2))
    return Scene("mixed", parts, 1.0e-4, dx, 2048, bodies, mats, substeps=100)
