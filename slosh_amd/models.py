"""Material, boundary-condition, shape and body constructors (reference API mirror).

Mirrors the reference's host-side constructors so scenes built here carry the
same bytes the reference would upload:

* ``ParticleModel`` variants and tags (src/solver/particle_model.rs:9-80, 87-102)
* ``ElasticCoefficients::from_young_modulus`` (src/models/mod.rs:52-59)
* ``DruckerPrager::new`` / ``from_lame`` defaults 35 deg / 9 deg / 0.2 / 10 deg
  (src/models/drucker_prager.rs:35-55) and the default plastic state {1, 1, 0}
  (src/models/drucker_prager.rs:85-92)
* ``GpuBoundaryCondition`` constants (src/solver/boundary_condition.rs:12-65)
* ``GpuShape::cuboid`` / ``ball`` / ``capsule`` / ``cone`` / ``cylinder`` (src/rbd/shapes.rs:82-201)
* ``ParticleDynamics::new``: V0 = (2r)^3, m = V0 * rho (src/solver/particle.rs:66-84)

All arithmetic is float32, as the reference computes these in f32.
"""
import numpy as np

from . import layouts as L

F32 = np.float32

ELASTIC_LINEAR = 0       # fixed-corotated, shaders/slosh/models/linear_elasticity.slang
ELASTIC_NEO_HOOKEAN = 1  # shaders/slosh/models/neo_hookean_elasticity.slang
SAND_LINEAR = 2          # Drucker-Prager + fixed-corotated
SAND_NEO_HOOKEAN = 3     # Drucker-Prager + neo-Hookean

STICK, SLIP, SEPARATE, FRICTION_Z_UP, NON_REFLECTING, DISABLED = range(6)

SHAPE_BALL, SHAPE_CUBOID, SHAPE_CAPSULE, SHAPE_CONE, SHAPE_CYLINDER, SHAPE_POLYLINE, SHAPE_TRIMESH = range(7)


def f32_bits(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32)


def lame_lambda_mu(young_modulus, poisson_ratio):
    """src/models/mod.rs:10-21 in float32."""
    e, nu = F32(young_modulus), F32(poisson_ratio)
    lam = (e * nu) / ((F32(1.0) + nu) * (F32(1.0) - F32(2.0) * nu))
    mu = e / (F32(2.0) * (F32(1.0) + nu))
    return F32(lam), F32(mu)


def to_radians(deg):
    """f32::to_radians: deg * (PI / 180)."""
    return F32(deg) * (F32(np.pi) / F32(180.0))


def elastic_coefficients(young_modulus, poisson_ratio):
    lam, mu = lame_lambda_mu(young_modulus, poisson_ratio)
    return (lam, mu, F32(0.5))


def _record(tag, floats):
    rec = np.zeros(13, dtype=np.uint32)
    rec[0] = tag
    fl = np.asarray(floats, dtype=np.float32)
    rec[1:1 + len(fl)] = fl.view(np.uint32)
    return rec


def elastic(young_modulus, poisson_ratio):
    """ParticleModel::elastic -> ElasticLinear (fixed-corotated)."""
    return _record(ELASTIC_LINEAR, elastic_coefficients(young_modulus, poisson_ratio))


def elastic_neo_hookean(young_modulus, poisson_ratio):
    return _record(ELASTIC_NEO_HOOKEAN, elastic_coefficients(young_modulus, poisson_ratio))


def elastic_from_lame(tag, lam, mu, cfl=0.5):
    """ElasticLinear / ElasticNeoHookean with explicit Lame parameters (public fields)."""
    return _record(tag, (F32(lam), F32(mu), F32(cfl)))


def drucker_prager(young_modulus, poisson_ratio, friction_angle_deg=35.0):
    if young_modulus > 0:
        lam, mu = lame_lambda_mu(young_modulus, poisson_ratio)
    else:
        lam, mu = F32(-1.0), F32(-1.0)
    return (to_radians(friction_angle_deg), to_radians(9.0), F32(0.2), to_radians(10.0), lam, mu)


def sand(young_modulus, poisson_ratio, friction_angle_deg=35.0, neo_hookean=False):
    """ParticleModel::sand (SandLinear) / sand_neo_hookean (SandNeoHookean)."""
    state = (F32(1.0), F32(1.0), F32(0.0))
    plastic = drucker_prager(young_modulus, poisson_ratio, friction_angle_deg)
    el = elastic_coefficients(young_modulus, poisson_ratio)
    return _record(SAND_NEO_HOOKEAN if neo_hookean else SAND_LINEAR, state + plastic + el)


def boundary_condition(ty, friction=0.0):
    bc = np.zeros(1, dtype=L.BOUNDARY_COND)
    bc["ty"] = ty
    bc["friction"] = friction
    return bc[0]


def cuboid(half_extents):
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0, :3] = half_extents
    sh["a"][0, 3] = np.uint32(SHAPE_CUBOID).view(np.float32)
    return sh[0]


def ball(radius):
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0, 0] = radius
    sh["a"][0, 3] = np.uint32(SHAPE_BALL).view(np.float32)
    return sh[0]


def capsule(a, b, radius):
    """GpuShape::capsule (src/rbd/shapes.rs:120-132): segment a..b, radius in b.w."""
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0, :3] = a
    sh["a"][0, 3] = np.uint32(SHAPE_CAPSULE).view(np.float32)
    sh["b"][0, :3] = b
    sh["b"][0, 3] = radius
    return sh[0]


def cone(half_height, radius):
    """GpuShape::cone (src/rbd/shapes.rs:181-187): axis y, apex at +half_height."""
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0, 0] = half_height
    sh["a"][0, 1] = radius
    sh["a"][0, 3] = np.uint32(SHAPE_CONE).view(np.float32)
    return sh[0]


def cylinder(half_height, radius):
    """GpuShape::cylinder (src/rbd/shapes.rs:195-201): axis y."""
    sh = np.zeros(1, dtype=L.SHAPE)
    sh["a"][0, 0] = half_height
    sh["a"][0, 1] = radius
    sh["a"][0, 3] = np.uint32(SHAPE_CYLINDER).view(np.float32)
    return sh[0]


def quat_from_scaled_axis(v):
    v = np.asarray(v, dtype=np.float64)
    ang = np.linalg.norm(v)
    if ang == 0.0:
        return np.array([0.0, 0.0, 0.0, 1.0], dtype=np.float32)
    ax = v / ang
    return np.array([*(ax * np.sin(ang / 2)), np.cos(ang / 2)], dtype=np.float32)


def body(shape, translation, rotation=(0.0, 0.0, 0.0, 1.0), linvel=(0, 0, 0), angvel=(0, 0, 0),
         mass=None, inv_inertia_diag=None):
    """BodyDesc. ``mass=None`` gives the zero mass properties the reference uses for
    fixed bodies and one-way coupling (src/rbd/dynamics/body.rs:204-229)."""
    b = np.zeros(1, dtype=L.BODY)
    b["shape"] = shape
    b["pose"]["rotation"] = rotation
    b["pose"]["translation_scale"][0, :3] = translation
    b["pose"]["translation_scale"][0, 3] = 1.0
    b["vel"]["linear"] = linvel
    b["vel"]["angular"] = angvel
    if mass is not None:
        inv_m = 1.0 / mass
        for key in ("local_mprops", "mprops"):
            b[key]["inv_mass"] = inv_m
            if inv_inertia_diag is not None:
                m = np.zeros((3, 4), np.float32)
                m[0, 0], m[1, 1], m[2, 2] = inv_inertia_diag
                b[key]["inv_inertia"] = m
    return b[0]


def particle_arrays(positions, radius, density, model, velocities=None):
    """Builds the reference's SoA particle buffers (SoAParticles::new,
    src/solver/particle.rs:401-429) for ``Particle::new(position, radius, density, model)``.

    ``model`` is one 13-word record or an (N, 13) array. Returns a dict with
    pos4, kin, def_grad, props, models.
    """
    positions = np.asarray(positions, dtype=np.float32)
    n = positions.shape[0]
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float32), (n,))
    density = np.broadcast_to(np.asarray(density, dtype=np.float32), (n,))
    init_volume = ((radius * F32(2.0)) * (radius * F32(2.0)) * (radius * F32(2.0))).astype(np.float32)
    pos4 = np.zeros((n, 4), dtype=np.float32)
    pos4[:, :3] = positions
    kin = np.zeros(n, dtype=L.KINEMATICS)
    kin["mass"] = init_volume * density
    kin["enabled"] = 1
    if velocities is not None:
        kin["velocity"] = velocities
    def_grad = np.zeros((n, 3, 4), dtype=np.float32)
    def_grad[:, 0, 0] = def_grad[:, 1, 1] = def_grad[:, 2, 2] = 1.0
    props = np.zeros(n, dtype=L.PROPERTIES)
    props["init_volume"] = init_volume
    props["init_radius"] = radius
    props["phase"] = 1.0
    model = np.asarray(model, dtype=np.uint32)
    models = np.broadcast_to(model, (n, 13)).copy() if model.ndim == 1 else model.copy()
    return {"pos4": pos4, "kin": kin, "def_grad": def_grad, "props": props, "models": models}


def concat_particles(parts):
    """Concatenates particle arrays, keeping each array's exact (padded) record dtype."""
    out = {}
    for k in parts[0]:
        first = parts[0][k]
        total = sum(p[k].shape[0] for p in parts)
        arr = np.zeros((total,) + first.shape[1:], dtype=first.dtype)
        o = 0
        for p in parts:
            arr[o:o + p[k].shape[0]] = p[k]
            o += p[k].shape[0]
        out[k] = arr
    return out
