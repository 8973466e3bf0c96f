"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/libslosh_oracle.so (built by oracle/build.sh). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module;
the product path (slosh_amd) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from slosh_amd import layouts as L

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libslosh_oracle.so")
_lib = None


def build():
    subprocess.check_call(["sh", os.path.join(_HERE, "build.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        l = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        l.so_create.restype = vp
        l.so_create.argtypes = [vp, C.c_size_t, vp, vp, vp, vp, vp, C.c_size_t, vp, vp, C.c_float, C.c_uint32, C.c_int]
        l.so_destroy.argtypes = [vp]
        l.so_step.argtypes = [vp]
        l.so_step.restype = C.c_int
        l.so_set_threads.argtypes = [C.c_int]
        l.so_set_p2g_gather.argtypes = [C.c_int]
        for name in ("so_read_positions", "so_read_kinematics", "so_read_def_grad", "so_read_models",
                     "so_read_blocks", "so_read_nodes"):
            getattr(l, name).argtypes = [vp, vp]
        l.so_read_bodies.argtypes = [vp, vp, vp]
        l.so_num_blocks.argtypes = [vp]
        l.so_num_blocks.restype = C.c_uint32
        l.so_svd3.argtypes = [vp, C.c_int, vp, vp, vp]
        l.so_model_update.argtypes = [vp, vp, vp, C.c_int]
        l.so_prefix_sum.argtypes = [vp, C.c_size_t]
        l.so_pack_key.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        l.so_pack_key.restype = C.c_uint32
        l.so_hash.argtypes = [C.c_uint32]
        l.so_hash.restype = C.c_uint32
        l.so_timestep_bound.argtypes = [vp]
        l.so_timestep_bound.restype = C.c_uint32
        l.so_particle_timestep_bound.argtypes = [vp, vp, vp, vp, C.c_float]
        l.so_particle_timestep_bound.restype = C.c_uint32
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Sequential restatement of MpmData + MpmPipeline::launch_step."""

    def __init__(self, scene, svd_mode=1, threads=0):
        l = lib()
        if threads:
            l.so_set_threads(threads)
        p = scene.particles
        self._keep = [np.ascontiguousarray(p[k]) for k in ("pos4", "kin", "def_grad", "props", "models")]
        self._params = np.array([scene.params], dtype=L.SIM_PARAMS)
        self._bodies = np.ascontiguousarray(scene.bodies)
        self._mats = np.ascontiguousarray(scene.materials)
        self.n = scene.n
        self.nb = len(self._bodies)
        self._s = l.so_create(_p(self._params), self.n, *[_p(a) for a in self._keep], self.nb,
                              _p(self._bodies), _p(self._mats), float(scene.cell_width),
                              scene.grid_capacity, svd_mode)
        if not self._s:
            raise MemoryError("oracle allocation failed")
        mv, mi = getattr(scene, "collision_vertices", None), getattr(scene, "collision_indices", None)
        if mv is not None and len(mv):
            self._mesh = (np.ascontiguousarray(mv, dtype=np.float32), np.ascontiguousarray(mi, dtype=np.uint32))
            l.so_set_collision_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
            l.so_set_collision_mesh(self._s, _p(self._mesh[0]), self._mesh[0].shape[0], _p(self._mesh[1]),
                                    self._mesh[1].size)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().so_destroy(self._s)
            self._s = None

    def step(self, k=1):
        for _ in range(k):
            rc = lib().so_step(self._s)
            if rc != 0:
                raise RuntimeError("oracle: hash map overflow")

    def positions(self):
        out = np.zeros((self.n, 4), np.float32)
        lib().so_read_positions(self._s, _p(out))
        return out

    def kinematics(self):
        out = np.zeros(self.n, L.KINEMATICS)
        lib().so_read_kinematics(self._s, _p(out))
        return out

    def def_grad(self):
        out = np.zeros((self.n, 3, 4), np.float32)
        lib().so_read_def_grad(self._s, _p(out))
        return out

    def models(self):
        out = np.zeros((self.n, 13), np.uint32)
        lib().so_read_models(self._s, _p(out))
        return out

    def bodies(self):
        poses = np.zeros(self.nb, L.SIM3)
        vels = np.zeros(self.nb, L.VELOCITY)
        if self.nb:
            lib().so_read_bodies(self._s, _p(poses), _p(vels))
        return poses, vels

    def timestep_bound_raw(self):
        return int(lib().so_timestep_bound(self._s))

    def timestep_bound(self):
        """Seconds, as GpuTimestepBounds::computed_dt (timestep_bound.rs:26-29)."""
        return float(np.float32(self.timestep_bound_raw()) / np.float32(1.0e12))

    def blocks(self):
        nb = lib().so_num_blocks(self._s)
        out = np.zeros(nb, L.BLOCK_INFO)
        if nb:
            lib().so_read_blocks(self._s, _p(out))
        return out

    def nodes(self):
        nb = lib().so_num_blocks(self._s)
        out = np.zeros((nb * 64, 4), np.float32)
        if nb:
            lib().so_read_nodes(self._s, _p(out))
        return out


def set_p2g_gather(gather):
    """P2G formulation of every oracle in the process: the shader's per-node gather
    (True) or the bit-identical block-local scatter (False, the default)."""
    lib().so_set_p2g_gather(1 if gather else 0)


def svd3(a, f64=True):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(9)
    u = np.zeros(9)
    s = np.zeros(3)
    v = np.zeros(9)
    lib().so_svd3(_p(a), 1 if f64 else 0, _p(u), _p(s), _p(v))
    return u.reshape(3, 3), s, v.reshape(3, 3)


def model_update(model, def_grad, svd_mode=1):
    m = np.array(model, dtype=np.uint32).copy()
    f = np.ascontiguousarray(def_grad, dtype=np.float32).reshape(9).copy()
    st = np.zeros(9, np.float32)
    lib().so_model_update(_p(m), _p(f), _p(st), svd_mode)
    return m, f.reshape(3, 3), st.reshape(3, 3)


def particle_timestep_bound(model, kin, def_grad, props, cell_width):
    """One particle's bound (u32 fixed point, 1e-12 s units)."""
    m = np.array(model, dtype=np.uint32).reshape(13).copy()
    k = np.array([kin], dtype=L.KINEMATICS)
    f = np.ascontiguousarray(def_grad, dtype=np.float32).reshape(1, 3, 4).copy()
    pr = np.array([props], dtype=L.PROPERTIES)
    return int(lib().so_particle_timestep_bound(_p(m), _p(k), _p(f), _p(pr), float(cell_width)))


def prefix_sum(v):
    a = np.ascontiguousarray(v, dtype=np.uint32).copy()
    lib().so_prefix_sum(_p(a), a.size)
    return a


def pack_key(x, y, z):
    return lib().so_pack_key(x, y, z)


def hash_key(k):
    return lib().so_hash(k)


class _Mesh(C.Structure):
    _fields_ = [("vtx", C.c_void_p), ("idx", C.c_void_p), ("nv", C.c_size_t), ("ni", C.c_size_t)]


def project_point_on_boundary(shape, pose, point, mesh=None, max_dist=1.0e30):
    """Shape::project_point_on_boundary (shape.slang:180-201) and the trimesh branch of
    collide (collide.slang:120-130): (handled, projection, inside) for one collider shape
    record (layouts.SHAPE) at one pose (layouts.SIM3). `mesh` = (collision vertices (n, 4)
    float32, collision indices u32) for trimesh shapes; `max_dist` bounds the trimesh query."""
    l = lib()
    l.so_project_point_on_boundary.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                               C.c_void_p, C.c_void_p]
    sh = np.array([shape], dtype=L.SHAPE)
    po = np.array([pose], dtype=L.SIM3)
    pt = np.ascontiguousarray(point, dtype=np.float32)
    out = np.zeros(3, np.float32)
    inside = C.c_int(0)
    m = None
    if mesh is not None:
        mv = np.ascontiguousarray(mesh[0], dtype=np.float32)
        mi = np.ascontiguousarray(mesh[1], dtype=np.uint32)
        m = _Mesh(mv.ctypes.data, mi.ctypes.data, mv.shape[0], mi.size)
    handled = l.so_project_point_on_boundary(_p(sh), _p(po), _p(pt), C.byref(m) if m is not None else None,
                                             float(max_dist), _p(out), C.byref(inside))
    return bool(handled), out, bool(inside.value)


def grid_update_node(bodies, mats, h, params, cell_pos, mvm):
    l = lib()
    l.so_grid_update_node.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    b = np.ascontiguousarray(np.asarray(bodies, dtype=L.BODY).reshape(-1))
    m = np.ascontiguousarray(np.asarray(mats, dtype=L.BOUNDARY_COND).reshape(-1))
    pr = np.array([params], dtype=L.SIM_PARAMS)
    cp = np.ascontiguousarray(cell_pos, dtype=np.float32)
    io = np.ascontiguousarray(mvm, dtype=np.float32).copy()
    nd = np.zeros(4, np.float32)
    l.so_grid_update_node(_p(b), _p(m), len(b), float(h), _p(pr), _p(cp), _p(io), _p(nd))
    return io, nd
