"""Python mirror of slosh's pipeline API over the C ABI (include/slosh_hip.h).

* ``MpmPipeline``  -- ``MpmPipeline::new`` / ``new_with_kernels`` + ``launch_step``
                      (src/pipeline.rs:456-770)
* ``MpmPipelineKernels`` -- kernel-family flags (src/pipeline.rs:39-93)
* ``MpmData``      -- ``MpmData::new`` (src/pipeline.rs:311-353) plus the readback,
                      append and removal entry points of ``GpuParticles``
                      (src/solver/particle.rs:505-602)
* ``MpmPipelineHooks`` -- the six callbacks of src/pipeline.rs:146-238

Every call goes through libslosh_hip.so; there is no host fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from . import layouts as L


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Stream:
    """A HIP stream owned by the library (non-blocking, default priority)."""

    def __init__(self):
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_stream_create(C.byref(h)))
        self.handle = h

    def synchronize(self):
        _lib.check(_lib.load().slosh_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            _lib.load().slosh_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_event_create(C.byref(h)))
        self.handle = h

    def record(self, stream):
        _lib.check(_lib.load().slosh_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, stop):
        ms = C.c_float()
        _lib.check(_lib.load().slosh_event_elapsed_ms(self.handle, stop.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            _lib.load().slosh_event_destroy(self.handle)
        except Exception:
            pass


class MpmPipelineKernels:
    def __init__(self, builtin_transfers=True, cdf=True, rigid_particles=True, node_reset=True, grid_update=True,
                 particles_update=True):
        self.builtin_transfers = builtin_transfers
        self.cdf = cdf
        self.rigid_particles = rigid_particles
        self.node_reset = node_reset
        self.grid_update = grid_update
        self.particles_update = particles_update

    @classmethod
    def hooks_only(cls):
        return cls(builtin_transfers=False, cdf=False, rigid_particles=False)

    def to_c(self):
        return _lib.Kernels(*(int(bool(getattr(self, f))) for f, _ in _lib.Kernels._fields_))


class MpmPipelineHooks:
    """Override any of these; return None/0 for success. ``run_p2g`` / ``run_g2p``
    return True when they handled the pass (the built-in pass is then skipped)."""

    def after_particle_sort(self, data, stream): pass
    def run_p2g(self, data, stream): return False
    def run_g2p(self, data, stream): return False
    def after_p2g(self, data, stream): pass
    def after_grid_update(self, data, stream): pass
    def after_g2p(self, data, stream): pass
    def after_particles_update(self, data, stream): pass

    def _to_c(self, data):
        def wrap(name):
            fn = getattr(self, name)
            if getattr(type(self), name) is getattr(MpmPipelineHooks, name):
                return _lib.HOOK_FN()
            def cb(_d, s, _state):
                try:
                    fn(data, s)
                    return 0
                except Exception:
                    return -5
            return _lib.HOOK_FN(cb)

        def wrap_replace(name):
            fn = getattr(self, name)
            if getattr(type(self), name) is getattr(MpmPipelineHooks, name):
                return _lib.REPLACE_FN()
            def cb(_d, s, _state, handled):
                try:
                    handled[0] = 1 if fn(data, s) else 0
                    return 0
                except Exception:
                    return -5
            return _lib.REPLACE_FN(cb)

        h = _lib.Hooks(wrap("after_particle_sort"), wrap_replace("run_p2g"), wrap_replace("run_g2p"),
                       wrap("after_p2g"), wrap("after_grid_update"), wrap("after_g2p"),
                       wrap("after_particles_update"))
        return h


class MpmPipeline:
    def __init__(self, kernels=None):
        kernels = kernels or MpmPipelineKernels()
        self._k = kernels.to_c()
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_pipeline_create(C.byref(self._k), C.byref(h)))
        self.handle = h

    def launch_step(self, data, stream=None, hooks=None, timestamps=None):
        """Enqueues one substep (MpmPipeline::launch_step); never synchronizes."""
        ch = hooks._to_c(data) if hooks is not None else None
        data._hooks_keepalive = ch
        _lib.check(_lib.load().slosh_launch_step(self.handle, data.handle, stream.handle if stream else None,
                                                 C.byref(ch) if ch is not None else None, None,
                                                 timestamps.handle if timestamps else None))

    def launch_steps_graph(self, data, count, stream=None):
        """Enqueues ``count`` substeps by replaying a captured HIP graph."""
        _lib.check(_lib.load().slosh_launch_steps_graph(self.handle, data.handle, stream.handle if stream else None,
                                                        int(count)))

    def __del__(self):
        try:
            _lib.load().slosh_pipeline_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """`steps` substeps of `data` captured once into a HIP graph and replayed with ``launch``.

    The graph keeps its data alive; if the data is closed first, ``launch`` raises."""

    def __init__(self, pipeline, data, steps):
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_graph_capture(pipeline.handle, data.handle, int(steps), C.byref(h)))
        self.handle = h
        self.steps = steps
        self._data = data

    def launch(self, stream=None):
        _lib.check(_lib.load().slosh_graph_launch(self.handle, stream.handle if stream else None))

    def __del__(self):
        try:
            _lib.load().slosh_graph_destroy(self.handle)
        except Exception:
            pass


class Timestamps:
    PASSES = ["grid_sort", "p2g", "grid_update", "g2p", "particles_update", "integrate_bodies"]

    def __init__(self):
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_timestamps_create(C.byref(h)))
        self.handle = h

    def read(self):
        ms = np.zeros(6, np.float64)
        cnt = np.zeros(6, np.uint32)
        _lib.check(_lib.load().slosh_timestamps_read(self.handle, _p(ms), _p(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.PASSES)}

    def reset(self):
        _lib.load().slosh_timestamps_reset(self.handle)

    def __del__(self):
        try:
            _lib.load().slosh_timestamps_destroy(self.handle)
        except Exception:
            pass


def _particles_struct(particles, scene_ids=None):
    keep = {k: np.ascontiguousarray(particles[k]) for k in ("pos4", "kin", "def_grad", "props", "models")}
    assert keep["kin"].dtype == L.KINEMATICS and keep["props"].dtype == L.PROPERTIES
    keep["models"] = keep["models"].astype(np.uint32, copy=False)
    if scene_ids is not None:
        # A non-null pointer marks the data as batched, also when it starts empty.
        sid = np.ascontiguousarray(scene_ids, dtype=np.uint32)
        keep["scene"] = sid if sid.size else np.zeros(1, np.uint32)
    n = keep["pos4"].shape[0]
    st = _lib.Particles(n, keep["pos4"].ctypes.data, keep["kin"].ctypes.data, keep["def_grad"].ctypes.data,
                        keep["props"].ctypes.data, keep["models"].ctypes.data,
                        keep["scene"].ctypes.data if scene_ids is not None else None)
    return st, keep


class MpmData:
    """Device-resident simulation state (MpmData::new)."""

    def __init__(self, pipeline, params, particles, bodies, materials, cell_width, grid_capacity,
                 scene_ids=None, particle_capacity=0, collision_mesh=None):
        st, keep = _particles_struct(particles, scene_ids)
        self._params = np.array([params], dtype=L.SIM_PARAMS)
        bodies = np.ascontiguousarray(np.asarray(bodies, dtype=L.BODY).reshape(-1))
        materials = np.ascontiguousarray(np.asarray(materials, dtype=L.BOUNDARY_COND).reshape(-1))
        h = C.c_void_p()
        _lib.check(_lib.load().slosh_data_create(pipeline.handle, _p(self._params), C.byref(st),
                                                 int(particle_capacity), _p(bodies) if len(bodies) else None,
                                                 len(bodies), _p(materials) if len(materials) else None,
                                                 float(cell_width), int(grid_capacity), C.byref(h)))
        self.handle = h
        self.nbodies = len(bodies)
        self.batched = scene_ids is not None
        self.cell_width = float(np.float32(cell_width))
        if collision_mesh is not None:
            self.set_collision_mesh(*collision_mesh)

    @property
    def params(self):
        """The SimulationParams record last set from the host (creation or set_sim_params)."""
        return self._params[0]

    @classmethod
    def from_scene(cls, pipeline, scene, **kw):
        if getattr(scene, "collision_vertices", None) is not None and "collision_mesh" not in kw:
            kw["collision_mesh"] = (scene.collision_vertices, scene.collision_indices)
        return cls(pipeline, scene.params, scene.particles, scene.bodies, scene.materials, scene.cell_width,
                   scene.grid_capacity, **kw)

    def set_collision_mesh(self, vertices4, indices):
        """slosh_data_set_collision_mesh: the trimesh collision buffers ((n, 4) float32
        x, y, z, pad; u32 indices) in convert_trimesh_to_gpu's layout."""
        self._mesh = (np.ascontiguousarray(vertices4, dtype=np.float32).reshape(-1, 4),
                      np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1))
        _lib.check(_lib.load().slosh_data_set_collision_mesh(self.handle, _p(self._mesh[0]), self._mesh[0].shape[0],
                                                             _p(self._mesh[1]), self._mesh[1].size))

    @classmethod
    def from_batch(cls, pipeline, batch, **kw):
        return cls(pipeline, batch.params, batch.particles, batch.bodies, batch.materials, batch.cell_width,
                   batch.grid_capacity, scene_ids=batch.scene_ids, **kw)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().slosh_data_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def close(self):
        self.__del__()

    @property
    def n(self):
        return int(_lib.load().slosh_num_particles(self.handle))

    def _s(self, stream):
        return stream.handle if stream else None

    def set_sim_params(self, params):
        self._params = np.array([params], dtype=L.SIM_PARAMS)
        _lib.check(_lib.load().slosh_set_sim_params(self.handle, _p(self._params)))

    def timestep_bound(self, stream=None):
        """(seconds, raw u32 fixed point) -- WgTimestepBounds::compute_bounds."""
        dt = C.c_float()
        raw = C.c_uint32()
        _lib.check(_lib.load().slosh_timestep_bound(self.handle, self._s(stream), C.byref(dt), C.byref(raw)))
        return dt.value, raw.value

    def timestep_bound_noalloc(self, out_raw_ptr, stream=None):
        """The raw bound into device memory at `out_raw_ptr` (one u32), asynchronous,
        allocating nothing (static per-GPU scratch; at most 33554432 particles)."""
        _lib.check(_lib.load().slosh_timestep_bound_noalloc(self.handle, self._s(stream), C.c_void_p(out_raw_ptr)))

    def adaptive_substeps(self, base_dt, min_substeps, max_substeps, stream=None):
        """Substep count for one frame of length base_dt (src_testbed/step.rs:100-147);
        sets the simulation dt to base_dt / count."""
        n = C.c_uint32()
        _lib.check(_lib.load().slosh_adaptive_substeps(self.handle, self._s(stream), float(base_dt),
                                                       int(min_substeps), int(max_substeps), C.byref(n)))
        return n.value

    def set_svd_fp64(self, enable):
        _lib.check(_lib.load().slosh_set_svd_fp64(self.handle, 1 if enable else 0))

    def check_errors(self, stream=None):
        _lib.check(_lib.load().slosh_check_errors(self.handle, self._s(stream)))

    def positions(self, stream=None):
        out = np.zeros((self.n, 4), np.float32)
        _lib.check(_lib.load().slosh_read_positions(self.handle, self._s(stream), _p(out)))
        return out

    def kinematics(self, stream=None):
        out = np.zeros(self.n, L.KINEMATICS)
        _lib.check(_lib.load().slosh_read_kinematics(self.handle, self._s(stream), _p(out)))
        return out

    def def_grad(self, stream=None):
        out = np.zeros((self.n, 3, 4), np.float32)
        _lib.check(_lib.load().slosh_read_def_grad(self.handle, self._s(stream), _p(out)))
        return out

    def models(self, stream=None):
        out = np.zeros((self.n, 13), np.uint32)
        _lib.check(_lib.load().slosh_read_models(self.handle, self._s(stream), _p(out)))
        return out

    def properties(self, stream=None):
        out = np.zeros(self.n, L.PROPERTIES)
        _lib.check(_lib.load().slosh_read_properties(self.handle, self._s(stream), _p(out)))
        return out

    def scene_ids(self, stream=None):
        out = np.zeros(self.n, np.uint32)
        _lib.check(_lib.load().slosh_read_scene_ids(self.handle, self._s(stream), _p(out)))
        return out

    def bodies(self, stream=None):
        poses = np.zeros(self.nbodies, L.SIM3)
        vels = np.zeros(self.nbodies, L.VELOCITY)
        _lib.check(_lib.load().slosh_read_bodies(self.handle, self._s(stream), _p(poses), _p(vels)))
        return poses, vels

    def num_blocks(self, stream=None):
        b = C.c_uint32()
        _lib.check(_lib.load().slosh_num_blocks(self.handle, self._s(stream), C.byref(b)))
        return b.value

    def blocks(self, stream=None):
        nb = self.num_blocks(stream)
        out = np.zeros(nb, L.BLOCK_INFO)
        w = C.c_uint32()
        if nb:
            _lib.check(_lib.load().slosh_read_blocks(self.handle, self._s(stream), _p(out), nb, C.byref(w)))
        return out[:w.value] if nb else out

    def block_scenes(self, stream=None):
        nb = self.num_blocks(stream)
        out = np.zeros(nb, np.uint32)
        if nb:
            _lib.check(_lib.load().slosh_read_block_scenes(self.handle, self._s(stream), _p(out), nb))
        return out

    def nodes(self, stream=None):
        nb = self.num_blocks(stream)
        out = np.zeros((nb * 64, 4), np.float32)
        if nb:
            _lib.check(_lib.load().slosh_read_nodes(self.handle, self._s(stream), _p(out), nb))
        return out

    def append(self, particles, scene_ids=None, stream=None):
        st, keep = _particles_struct(particles, scene_ids)
        _lib.check(_lib.load().slosh_append_particles(self.handle, self._s(stream), C.byref(st)))

    def reserve_staging(self, nbytes):
        """Pins `nbytes` of host memory as the ring appends stage in (200 bytes per
        particle); 0 releases it."""
        _lib.check(_lib.load().slosh_reserve_staging(self.handle, int(nbytes)))

    def shift_remove(self, start, end, stream=None):
        _lib.check(_lib.load().slosh_remove_particles(self.handle, self._s(stream), int(start), int(end)))

    def remove_scene(self, scene_id, stream=None):
        r = C.c_size_t()
        _lib.check(_lib.load().slosh_remove_scene(self.handle, self._s(stream), int(scene_id), C.byref(r)))
        return r.value

    @property
    def tuning(self):
        t = _lib.Tuning()
        _lib.check(_lib.load().slosh_get_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = self.tuning
        t.update(kw)
        _lib.check(_lib.load().slosh_set_tuning(self.handle, C.byref(_lib.Tuning(**t))))

    def autotune(self, pipeline, stream=None, trial_steps=10):
        """Picks the fastest launch shape on a private copy of this state (the
        state itself does not advance); returns it as a dict."""
        t = _lib.Tuning()
        _lib.check(_lib.load().slosh_autotune(pipeline.handle, self.handle, self._s(stream), int(trial_steps),
                                              C.byref(t)))
        return t.as_dict()

    def remove_scenes(self, scene_ids, stream=None):
        ids = np.ascontiguousarray(scene_ids, dtype=np.uint32)
        r = C.c_size_t()
        _lib.check(_lib.load().slosh_remove_scenes(self.handle, self._s(stream), _p(ids), ids.size, C.byref(r)))
        return r.value

    def scene_size(self, scene_id):
        return int(_lib.load().slosh_scene_size(self.handle, int(scene_id)))

    def read_scene(self, scene_id, stream=None, positions_only=False):
        """(positions [m, 4], kinematics [m], def_grad [m, 3, 4]) of one scene in creation
        order, in the formats of positions() / kinematics() / def_grad(); with
        positions_only, just the positions."""
        m = self.scene_size(scene_id)
        pos = np.empty((m, 4), np.float32)
        kin = None if positions_only else np.zeros(m, L.KINEMATICS)
        f = None if positions_only else np.zeros((m, 3, 4), np.float32)
        cnt = C.c_size_t()
        _lib.check(_lib.load().slosh_read_scene(self.handle, self._s(stream), int(scene_id), _p(pos),
                                                _p(kin) if kin is not None else None,
                                                _p(f) if f is not None else None, m, C.byref(cnt)))
        return pos if positions_only else (pos, kin, f)

    def write_kinematics(self, kin, stream=None):
        kin = np.ascontiguousarray(kin, dtype=L.KINEMATICS)
        _lib.check(_lib.load().slosh_write_kinematics(self.handle, self._s(stream), _p(kin)))

    def device_view(self):
        v = _lib.DeviceView()
        _lib.check(_lib.load().slosh_device_view_get(self.handle, C.byref(v)))
        return v


def device_count():
    n = C.c_int(0)
    rc = _lib.load().slosh_device_count(C.byref(n))
    return n.value if rc == 0 else 0
