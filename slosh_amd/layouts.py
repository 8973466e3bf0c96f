"""Host-side byte layouts of the reference's particle, body and grid records.

Every dtype here mirrors one of slosh's `#[repr(C)]` / encase records so that
buffers handed across the C ABI (include/slosh_hip.h) have the reference's
bytes:

* ``POSITION``      -- ``ParticlePosition`` = glam ``Vec4`` (src/solver/particle.rs:388-389)
* ``KINEMATICS``    -- ``Kinematics``, 96 B (src/solver/particle.rs:132-147)
* ``MATRIX``        -- ``Matrix`` as 3 x vec4 (48 B); row r of the shader matrix
                       is glam column r (shaders/slosh/solver/particle.slang:16)
* ``PROPERTIES``    -- ``ParticleProperties``, 20 B (src/solver/particle.rs:154-167)
* ``MODEL``         -- ``GpuParticleModel``, 52 B: u32 tag + payload
                       (src/solver/particle_model.rs:87-102)
* ``BOUNDARY_COND`` -- ``GpuBoundaryCondition`` (src/solver/boundary_condition.rs:6-10)
* ``SHAPE``         -- ``GpuShape`` 2 x vec4 (src/rbd/shapes.rs:66-70)
* ``SIM3``          -- ``GpuSim3`` pose: quaternion xyzw, translation xyz + scale
* ``VELOCITY``      -- ``GpuVelocity`` (src/rbd/dynamics/body.rs:24-32)
* ``MASS_PROPS``    -- ``GpuMassProperties`` (src/rbd/dynamics/body.rs:34-44)
* ``BODY``          -- ``BodyDesc`` (src/rbd/dynamics/body.rs:93-106)
* ``SIM_PARAMS``    -- ``SimulationParams`` (src/solver/params.rs:13-21)
* ``BLOCK_INFO``    -- per active block: virtual id, first particle, counts and
                       the 10 within-block sort-bucket counts
                       (``GpuActiveBlockHeader``, src/grid/grid.rs:335-346)
"""
import numpy as np

POSITION = np.dtype((np.float32, 4))
MATRIX = np.dtype((np.float32, (3, 4)))

KINEMATICS = np.dtype({
    "names": ["affine", "velocity", "force_dt", "vel_grad_det", "mass", "enabled"],
    "formats": [MATRIX, (np.float32, 3), (np.float32, 3), np.float32, np.float32, np.uint32],
    "offsets": [0, 48, 64, 76, 80, 84],
    "itemsize": 96,
})

PROPERTIES = np.dtype({
    "names": ["init_volume", "init_radius", "damping", "phase", "fixed"],
    "formats": [np.float32, np.float32, np.float32, np.float32, np.uint32],
    "offsets": [0, 4, 8, 12, 16],
    "itemsize": 20,
})

MODEL = np.dtype((np.uint32, 13))

BOUNDARY_COND = np.dtype({"names": ["ty", "friction"], "formats": [np.uint32, np.float32],
                          "offsets": [0, 4], "itemsize": 8})

SHAPE = np.dtype({"names": ["a", "b"], "formats": [(np.float32, 4), (np.float32, 4)],
                  "offsets": [0, 16], "itemsize": 32})

SIM3 = np.dtype({"names": ["rotation", "translation_scale"], "formats": [(np.float32, 4), (np.float32, 4)],
                 "offsets": [0, 16], "itemsize": 32})

VELOCITY = np.dtype({"names": ["linear", "angular"], "formats": [(np.float32, 3), (np.float32, 3)],
                     "offsets": [0, 16], "itemsize": 32})

MASS_PROPS = np.dtype({"names": ["inv_inertia", "inv_mass", "com"],
                       "formats": [MATRIX, (np.float32, 3), (np.float32, 3)],
                       "offsets": [0, 48, 64], "itemsize": 80})

BODY = np.dtype({"names": ["local_mprops", "mprops", "vel", "pose", "shape"],
                 "formats": [MASS_PROPS, MASS_PROPS, VELOCITY, SIM3, SHAPE],
                 "offsets": [0, 80, 160, 192, 224], "itemsize": 256})

SIM_PARAMS = np.dtype({"names": ["gravity", "dt"], "formats": [(np.float32, 3), np.float32],
                       "offsets": [0, 12], "itemsize": 16})

BLOCK_INFO = np.dtype({
    "names": ["vid", "first_particle", "num_particles", "num_particles_with_extras", "bucket_counts"],
    "formats": [(np.int32, 3), np.uint32, np.uint32, np.uint32, (np.uint32, 10)],
    "offsets": [0, 12, 16, 20, 24],
    "itemsize": 64,
})

assert KINEMATICS.itemsize == 96 and PROPERTIES.itemsize == 20 and MODEL.itemsize == 52
assert BODY.itemsize == 256 and MATRIX.itemsize == 48


def mat_to_slang(m):
    """(..., 3, 3) row-major shader matrices -> (..., 3, 4) padded record rows."""
    m = np.asarray(m, dtype=np.float32)
    out = np.zeros(m.shape[:-2] + (3, 4), dtype=np.float32)
    out[..., :3] = m
    return out


def mat_from_slang(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.float32)[..., :3])
