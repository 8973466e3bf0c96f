"""ctypes binding of libslosh_hip.so (declarations of include/slosh_hip.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslosh_hip.so")

SLOSH_OK = 0
ERRORS = {-1: "SLOSH_E_INVALID", -2: "SLOSH_E_HIP", -3: "SLOSH_E_NO_MEMORY", -4: "SLOSH_E_GRID_OVERFLOW",
          -5: "SLOSH_E_HOOK", -6: "SLOSH_E_UNSUPPORTED", -7: "SLOSH_E_NO_KERNELS", -8: "SLOSH_E_KEY_RANGE"}

# Every symbol include/slosh_hip.h declares.
EXPORTS = [
    "slosh_pipeline_create", "slosh_pipeline_destroy", "slosh_kernels_default", "slosh_data_create",
    "slosh_data_destroy", "slosh_data_set_collision_mesh", "slosh_launch_step", "slosh_launch_steps_graph",
    "slosh_set_svd_fp64", "slosh_set_sim_params",
    "slosh_check_errors", "slosh_num_particles", "slosh_read_positions", "slosh_read_kinematics",
    "slosh_read_def_grad", "slosh_read_models", "slosh_read_properties", "slosh_read_scene_ids",
    "slosh_read_bodies", "slosh_num_blocks", "slosh_read_blocks", "slosh_read_nodes", "slosh_read_block_scenes",
    "slosh_append_particles", "slosh_remove_particles", "slosh_remove_scene", "slosh_remove_scenes",
    "slosh_scene_size", "slosh_read_scene", "slosh_write_kinematics",
    "slosh_device_view_get", "slosh_timestamps_create", "slosh_timestamps_destroy", "slosh_timestamps_read",
    "slosh_timestamps_reset", "slosh_stream_create", "slosh_stream_destroy", "slosh_stream_synchronize",
    "slosh_event_create", "slosh_event_destroy", "slosh_event_record", "slosh_event_elapsed_ms",
    "slosh_device_count", "slosh_last_error", "slosh_version", "slosh_graph_capture", "slosh_graph_launch",
    "slosh_graph_destroy", "slosh_set_tuning", "slosh_get_tuning", "slosh_autotune", "slosh_timestep_bound",
    "slosh_adaptive_substeps", "slosh_timestep_bound_noalloc", "slosh_reserve_staging",
]


class SloshError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Tuning(C.Structure):
    _fields_ = [("p2g_chunk", C.c_uint32), ("g2p_threads", C.c_uint32), ("block_wgs", C.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Kernels(C.Structure):
    _fields_ = [("builtin_transfers", C.c_uint32), ("cdf", C.c_uint32), ("rigid_particles", C.c_uint32),
                ("node_reset", C.c_uint32), ("grid_update", C.c_uint32), ("particles_update", C.c_uint32)]


class Particles(C.Structure):
    _fields_ = [("len", C.c_size_t), ("positions", C.c_void_p), ("kinematics", C.c_void_p),
                ("def_grad", C.c_void_p), ("properties", C.c_void_p), ("models", C.c_void_p),
                ("scene_ids", C.c_void_p)]


HOOK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
REPLACE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int))


class Hooks(C.Structure):
    _fields_ = [("after_particle_sort", HOOK_FN), ("run_p2g", REPLACE_FN), ("run_g2p", REPLACE_FN),
                ("after_p2g", HOOK_FN), ("after_grid_update", HOOK_FN), ("after_g2p", HOOK_FN),
                ("after_particles_update", HOOK_FN)]


class DeviceView(C.Structure):
    _fields_ = [("num_particles", C.c_size_t), ("particle_capacity", C.c_size_t), ("grid_capacity", C.c_uint32),
                ("cell_width", C.c_float), ("positions", C.c_void_p), ("vel_mass", C.c_void_p),
                ("affine", C.c_void_p), ("def_grad", C.c_void_p), ("vel_grad_det", C.c_void_p),
                ("particle_ids", C.c_void_p), ("force_dt", C.c_void_p), ("nodes", C.c_void_p),
                ("sorted_ids", C.c_void_p), ("counters", C.c_void_p),
                ("block_list", C.c_void_p)]


_lib = None


def load():
    """Loads the in-tree library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m slosh_amd.build` (the HIP library is required)")
    l = C.CDLL(LIB_PATH)
    vp, sz, u32, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
    sig = {
        "slosh_pipeline_create": ([vp, vp], i32),
        "slosh_pipeline_destroy": ([vp], None),
        "slosh_kernels_default": ([vp], None),
        "slosh_data_create": ([vp, vp, vp, sz, vp, sz, vp, C.c_float, u32, vp], i32),
        "slosh_data_destroy": ([vp], None),
        "slosh_launch_step": ([vp, vp, vp, vp, vp, vp], i32),
        "slosh_launch_steps_graph": ([vp, vp, vp, u32], i32),
        "slosh_set_sim_params": ([vp, vp], i32),
        "slosh_set_svd_fp64": ([vp, i32], i32),
        "slosh_data_set_collision_mesh": ([vp, vp, C.c_size_t, vp, C.c_size_t], i32),
        "slosh_check_errors": ([vp, vp], i32),
        "slosh_num_particles": ([vp], sz),
        "slosh_read_positions": ([vp, vp, vp], i32),
        "slosh_read_kinematics": ([vp, vp, vp], i32),
        "slosh_read_def_grad": ([vp, vp, vp], i32),
        "slosh_read_models": ([vp, vp, vp], i32),
        "slosh_read_properties": ([vp, vp, vp], i32),
        "slosh_read_scene_ids": ([vp, vp, vp], i32),
        "slosh_read_bodies": ([vp, vp, vp, vp], i32),
        "slosh_num_blocks": ([vp, vp, vp], i32),
        "slosh_read_blocks": ([vp, vp, vp, u32, vp], i32),
        "slosh_read_nodes": ([vp, vp, vp, u32], i32),
        "slosh_read_block_scenes": ([vp, vp, vp, u32], i32),
        "slosh_append_particles": ([vp, vp, vp], i32),
        "slosh_remove_particles": ([vp, vp, sz, sz], i32),
        "slosh_remove_scene": ([vp, vp, u32, vp], i32),
        "slosh_remove_scenes": ([vp, vp, vp, sz, vp], i32),
        "slosh_set_tuning": ([vp, vp], i32),
        "slosh_get_tuning": ([vp, vp], i32),
        "slosh_autotune": ([vp, vp, vp, u32, vp], i32),
        "slosh_timestep_bound": ([vp, vp, vp, vp], i32),
        "slosh_timestep_bound_noalloc": ([vp, vp, vp], i32),
        "slosh_reserve_staging": ([vp, sz], i32),
        "slosh_adaptive_substeps": ([vp, vp, C.c_float, u32, u32, vp], i32),
        "slosh_scene_size": ([vp, u32], sz),
        "slosh_read_scene": ([vp, vp, u32, vp, vp, vp, sz, vp], i32),
        "slosh_write_kinematics": ([vp, vp, vp], i32),
        "slosh_device_view_get": ([vp, vp], i32),
        "slosh_timestamps_create": ([vp], i32),
        "slosh_timestamps_destroy": ([vp], None),
        "slosh_timestamps_read": ([vp, vp, vp], i32),
        "slosh_timestamps_reset": ([vp], None),
        "slosh_stream_create": ([vp], i32),
        "slosh_stream_destroy": ([vp], i32),
        "slosh_stream_synchronize": ([vp], i32),
        "slosh_event_create": ([vp], i32),
        "slosh_event_destroy": ([vp], i32),
        "slosh_event_record": ([vp, vp], i32),
        "slosh_event_elapsed_ms": ([vp, vp, vp], i32),
        "slosh_device_count": ([vp], i32),
        "slosh_graph_capture": ([vp, vp, u32, vp], i32),
        "slosh_graph_launch": ([vp, vp], i32),
        "slosh_graph_destroy": ([vp], None),
        "slosh_last_error": ([], C.c_char_p),
        "slosh_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(l, name)
        f.argtypes = args
        f.restype = res
    _lib = l
    return l


def check(rc):
    if rc != SLOSH_OK:
        raise SloshError(rc, load().slosh_last_error().decode(errors="replace"))
    return rc
