"""slosh_timestep_bound_noalloc against slosh_timestep_bound through the C ABI.

The no-allocation entry point keeps its partial minima in a static per-GPU
array that keeps the last call's contents; its result must equal the existing
entry point's bit for bit, also when a smaller call follows a larger one (the
stale partials of the larger call must not be read).
"""
import ctypes as C

import numpy as np
import pytest

from slosh_amd import pipeline as P
from slosh_amd import scenes

pytestmark = pytest.mark.gpu


class DeviceWord:
    """One u32 of device memory, allocated by the test (the caller), not by the entry point."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(4)) == 0

    def fill(self, value):
        assert self.hip.hipDeviceSynchronize() == 0
        host = C.c_uint32(value)
        assert self.hip.hipMemcpy(self.ptr, C.byref(host), C.c_size_t(4), 1) == 0  # host -> device

    def read(self, stream):
        stream.synchronize()
        host = C.c_uint32(0)
        assert self.hip.hipMemcpy(C.byref(host), self.ptr, C.c_size_t(4), 2) == 0  # device -> host
        return host.value

    def close(self):
        self.hip.hipFree(self.ptr)


def _bounds(data, st, out):
    """(existing entry point's raw bound, the no-allocation entry point's)."""
    _, raw = data.timestep_bound(st)
    out.fill(0x12345678)  # a value neither entry point returns here
    data.timestep_bound_noalloc(out.ptr.value, st)
    return raw, out.read(st)


def test_noalloc_bound_equals_existing_entry_point_large_then_small(gpu_pipeline):
    pl, st = gpu_pipeline
    out = DeviceWord()
    try:
        # First call: the 1M-particle dam break after a few substeps (489 partials).
        big = scenes.dam_break()
        scenes.stir(big, 77)
        d_big = P.MpmData.from_scene(pl, big)
        for _ in range(3):
            pl.launch_step(d_big, st)
        d_big.check_errors(st)
        print(f"\n[noalloc] call 1: n={big.n}", flush=True)
        ref, got = _bounds(d_big, st, out)
        assert ref != 0xFFFFFFFF and got == ref, (hex(ref), hex(got))

        # Second call, smaller input: one partial; the stale ones above it stay unread.
        small = scenes.mixed_test_scene(n_side=12)
        d_small = P.MpmData.from_scene(pl, small)
        pl.launch_step(d_small, st)
        d_small.check_errors(st)
        print(f"[noalloc] call 2: n={small.n}", flush=True)
        ref2, got2 = _bounds(d_small, st, out)
        assert got2 == ref2, (hex(ref2), hex(got2))
        # The smaller scene's bound differs from the larger one's, so reading a
        # stale partial would have shown.
        assert ref2 != ref

        # And again on the larger data, after more substeps (contents change).
        pl.launch_step(d_big, st)
        ref3, got3 = _bounds(d_big, st, out)
        assert got3 == ref3
        d_big.close()
        d_small.close()
    finally:
        out.close()


def test_noalloc_bound_ragged_sizes(gpu_pipeline):
    """Particle counts around the 2048-particle partial boundary, each after a
    larger call, against the existing entry point."""
    pl, st = gpu_pipeline
    out = DeviceWord()
    try:
        for n_side in (16, 13, 12, 10, 3):  # 4096, 2197, 1728, 1000, 27 particles
            sc = scenes.elastic_cube(n_side=n_side)
            scenes.stir(sc, 5 + n_side)
            d = P.MpmData.from_scene(pl, sc)
            pl.launch_step(d, st)
            ref, got = _bounds(d, st, out)
            print(f"[noalloc] n={sc.n} raw={ref}", flush=True)
            assert got == ref, (sc.n, hex(ref), hex(got))
            d.close()
    finally:
        out.close()
