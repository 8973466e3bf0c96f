"""Captured substeps stay consistent with the data they were captured from.

A captured substep carries step parameters in its kernel arguments (uniform mass,
collision-cache switch, buffers). Every change to those retires the captured
graphs: cached substep pairs of ``launch_steps_graph`` are recaptured, user
graphs refuse to launch. A user graph whose data was destroyed refuses to launch
instead of replaying over freed buffers.
"""
import copy
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from slosh_amd import _lib
from slosh_amd import pipeline as P
from slosh_amd import scenes
from tests.parity import state_errors, tol_for

pytestmark = pytest.mark.gpu


def _restart_scene(sc, data, st):
    rs = copy.copy(sc)
    pos = data.positions(st)
    pos[:, 3] = 0.0
    rs.particles = {"pos4": pos, "kin": data.kinematics(st), "def_grad": data.def_grad(st),
                    "props": data.properties(st), "models": data.models(st)}
    return rs


def _check(tag, data, orc, st, x0, steps):
    errs = state_errors((data.positions(st), data.kinematics(st), data.def_grad(st)),
                        (orc.positions(), orc.kinematics(), orc.def_grad()), x0)
    print(f"\n[{tag}] " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()))
    assert errs["enabled_diff"] == 0
    for k in ("x", "v", "affine", "F"):
        assert errs[k] <= tol_for(steps), (k, errs[k])
    assert np.array_equal(data.kinematics(st)["mass"], orc.kinematics()["mass"])


def test_graph_replay_after_written_masses(gpu_pipeline):
    """Pairs captured while every particle had one mass (read from the kernel
    arguments) must not be replayed after the masses were rewritten."""
    pl, st = gpu_pipeline
    sc = scenes.stir(scenes.elastic_cube(n_side=12), 21)
    data = P.MpmData.from_scene(pl, sc)
    pl.launch_step(data, st)
    pl.launch_steps_graph(data, 4, st)  # captures and replays the pair under the uniform mass
    g = P.Graph(pl, data, 2)
    g.launch(st)
    kin = data.kinematics(st)
    kin["mass"][::3] *= 2.0             # no force: a plain rewrite of the masses
    data.write_kinematics(kin, st)
    rs = _restart_scene(sc, data, st)
    assert np.array_equal(rs.particles["kin"]["mass"], kin["mass"])
    with pytest.raises(_lib.SloshError):
        g.launch(st)                    # captured before the rewrite: refused
    pl.launch_steps_graph(data, 1, st)  # re-primes the table eagerly
    pl.launch_steps_graph(data, 4, st)  # pairs recaptured with the written masses
    data.check_errors(st)
    orc = O.Oracle(rs)
    orc.step(5)
    _check("graph after written masses", data, orc, st, rs.particles["pos4"], 5)


class _DoubleMasses(P.MpmPipelineHooks):
    """after_particles_update: doubles the mass of every third storage slot
    through the device view (the way a user kernel would)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipStreamSynchronize.argtypes = [C.c_void_p]

    def after_particles_update(self, data, stream):
        v = data.device_view()
        n = int(v.num_particles)
        assert self.hip.hipStreamSynchronize(stream) == 0
        buf = np.empty((n, 4), np.float32)
        assert self.hip.hipMemcpy(buf.ctypes.data, v.vel_mass, 16 * n, 2) == 0  # device to host
        buf[::3, 3] *= 2.0
        assert self.hip.hipMemcpy(v.vel_mass, buf.ctypes.data, 16 * n, 1) == 0  # host to device


def test_graph_replay_after_hooked_step(gpu_pipeline):
    """A hooked substep may write masses and switches on the collision cache; the
    pairs captured before it must be recaptured (the fused G2P would otherwise
    write the captured uniform mass over the written ones)."""
    pl, st = gpu_pipeline
    sc = scenes.stir(scenes.elastic_cube(n_side=12), 22)  # static walls: no collision cache at first
    data = P.MpmData.from_scene(pl, sc)
    pl.launch_step(data, st)
    pl.launch_steps_graph(data, 4, st)
    pl.launch_step(data, st, hooks=_DoubleMasses())
    st.synchronize()
    rs = _restart_scene(sc, data, st)
    m = rs.particles["kin"]["mass"]
    assert len(np.unique(m[m > 0])) == 2  # the hook's writes are in place
    pl.launch_steps_graph(data, 1, st)
    pl.launch_steps_graph(data, 4, st)
    data.check_errors(st)
    orc = O.Oracle(rs)
    orc.step(5)
    _check("graph after hooked step", data, orc, st, rs.particles["pos4"], 5)


def test_user_graph_refused_after_mutation_or_destroyed_data(gpu_pipeline):
    pl, st = gpu_pipeline
    sc = scenes.elastic_cube(n_side=8)
    data = P.MpmData.from_scene(pl, sc)
    pl.launch_step(data, st)
    g = P.Graph(pl, data, 2)
    g.launch(st)
    g.launch(st)
    data.set_sim_params(data.params)  # any mutation retires the graph
    with pytest.raises(_lib.SloshError):
        g.launch(st)
    pl.launch_step(data, st)
    g2 = P.Graph(pl, data, 2)
    g2.launch(st)
    st.synchronize()
    data.close()                      # the data goes first: the graph is detached
    with pytest.raises(_lib.SloshError):
        g2.launch(st)
    del g2, g                         # destroying detached graphs is safe
    # A graph object alone keeps its data alive.
    d3 = P.MpmData.from_scene(pl, sc)
    pl.launch_step(d3, st)
    g3 = P.Graph(pl, d3, 2)
    del d3
    g3.launch(st)
    st.synchronize()
