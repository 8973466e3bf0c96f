"""C-ABI checks that need no GPU: the library loads and exports every symbol the
public header declares; host-side argument validation returns error codes."""
import ctypes as C
import os
import re

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
