"""C-ABI checks that need no GPU: the library loads and exports every symbol the
public header declares; host-side argument validation returns error codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from slosh_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "slosh_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(slosh_[a-z0-9_]+)\s*\(", text))
    # typedef'd function-pointer types are not exports
    names -= set(re.findall(r"\(\*\s*(slosh_[a-z0-9_]+)\)", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert len(declared) >= 35
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def test_version_and_error_strings():
    lib = _lib.load()
    assert b"gfx950" in lib.slosh_version()


def test_invalid_arguments_return_codes_without_gpu_work():
    lib = _lib.load()
    assert lib.slosh_pipeline_create(None, None) == -1
    assert b"null" in lib.slosh_last_error()
    assert lib.slosh_data_create(None, None, None, 0, None, 0, None, C.c_float(1.0), 16, None) == -1
    assert lib.slosh_launch_step(None, None, None, None, None, None) == -1
    assert lib.slosh_set_sim_params(None, None) == -1


def test_struct_sizes_match_reference_layouts():
    from slosh_amd import layouts as L
    assert L.KINEMATICS.itemsize == 96      # Kinematics
    assert L.PROPERTIES.itemsize == 20      # ParticleProperties
    assert L.MODEL.itemsize == 52           # GpuParticleModel
    assert L.BODY.itemsize == 256           # BodyDesc = 2 x 80 + 32 + 32 + 32
    assert L.BLOCK_INFO.itemsize == 64
    assert C.sizeof(_lib.Particles) == 56
    assert C.sizeof(_lib.Kernels) == 24


def _particles(pos, h=1.0 / 128):
    from slosh_amd import models as M
    return M.particle_arrays(np.asarray(pos, np.float32), np.float32(h / 4), 1000.0, M.elastic_neo_hookean(1e5, 0.3))


def test_block_key_range_is_checked_before_any_gpu_work():
    """Creation rejects particles whose block (or its "+1" neighbour) has no exact
    key: batched blocks must lie in [-16, 46] per axis (6 bits, offset 16),
    single-scene blocks in x, z [-1023, 1023], y [-511, 511]. The check runs on
    the host before any device allocation, so it needs no GPU."""
    from slosh_amd import layouts as L
    from slosh_amd import pipeline as P
    h = 1.0 / 128
    pl = P.MpmPipeline()

    def create(pos, scene_ids):
        parts = _particles(pos, h)
        st, keep = P._particles_struct(parts, scene_ids)
        prm = np.zeros(1, dtype=L.SIM_PARAMS)
        prm["dt"] = 1e-4
        out = C.c_void_p()
        rc = _lib.load().slosh_data_create(pl.handle, C.c_void_p(prm.ctypes.data), C.byref(st), 0, None, 0, None,
                                           C.c_float(h), 1024, C.byref(out))
        return rc, _lib.load().slosh_last_error()

    ids = np.zeros(2, np.uint32)
    # x = 1.48: round(x / h) - 1 = 188, block 47 -> out of the batched range
    rc, msg = create([[0.5, 0.5, 0.5], [1.48, 0.5, 0.5]], ids)
    assert rc == -8 and b"particle 1" in msg and b"[-16, 46]" in msg
    # y = -0.52: block -18
    rc, _ = create([[0.5, -0.52, 0.5], [0.5, 0.5, 0.5]], ids)
    assert rc == -8
    # single scene: y blocks must stay >= -511; y = -16.0 is block -513
    rc, _ = create([[0.5, -16.0, 0.5]], None)
    assert rc == -8
    rc, _ = create([[0.5, float("nan"), 0.5]], None)
    assert rc == -8
