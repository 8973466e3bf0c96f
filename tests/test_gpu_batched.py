"""Batched scenes (many independent scenes in one launch) on the GPU.

Each scene of a batch must give what the oracle gives for that scene alone: the
active-block set and per-block counts bit-exact (blocks compared per scene, keyed
by the scene-local virtual id) and x, v, affine, F within relative L2 <= 1e-5
after one substep and <= 1e-3 after 100. Device-side removal must keep the other
scenes' state bit for bit.
"""
import copy

import numpy as np
import pytest

from slosh_amd import pipeline as P
from slosh_amd import scenes
from oracle import oracle as O
from tests.parity import assert_blocks_equal, state_errors, tol_for

pytestmark = pytest.mark.gpu

TOL = {1: 1.0e-5, 100: 1.0e-3}


def small_batch():
    # Scenes 0..3 of the batched configuration at reduced sizes: neo-Hookean,
    # fixed-corotated and Drucker-Prager all occur among them.
    # Stirred (nonuniform initial velocity) so that APIC matrices are not round-off.
    counts = (3000, 1200, 5000, 2000)
    scs = [scenes.stir(scenes.batched_scene(i, n_override=n), 77 + i) for i, n in enumerate(counts)]
    tags = {int(sc.particles["models"][0][0]) for sc in scs}
    assert len(tags) >= 2
    return scenes.Batch(scs, ids=[0, 7, 3, 12])


@pytest.mark.parametrize("steps", [1, 100])
def test_batch_matches_each_scene_alone(gpu_pipeline, steps):
    pl, st = gpu_pipeline
    b = small_batch()
    data = P.MpmData.from_batch(pl, b)
    for _ in range(steps):
        pl.launch_step(data, st)
    data.check_errors(st)
    blocks, bscene = data.blocks(st), data.block_scenes(st)
    for sc, sid in zip(b.scenes, b.ids):
        orc = O.Oracle(sc, svd_mode=1)
        orc.step(steps)
        assert_blocks_equal(blocks[bscene == sid], orc.blocks())
        pos, kin, f = data.read_scene(sid, st)
        errs = state_errors((pos, kin, f), (orc.positions(), orc.kinematics(), orc.def_grad()),
                            sc.particles["pos4"])
        print(f"\n[batched scene {sid} (n={sc.n}) after {steps}] " +
              " ".join(f"{k}={v:.3e}" for k, v in errs.items() if k != "enabled_diff"))
        assert errs["enabled_diff"] == 0
        for k in ("x", "v", "affine", "F"):
            assert errs[k] <= TOL[steps], f"scene {sid} {k}: {errs[k]:.3e}"


def test_read_scene_matches_full_readback(gpu_pipeline):
    pl, st = gpu_pipeline
    b = small_batch()
    data = P.MpmData.from_batch(pl, b)
    for _ in range(3):
        pl.launch_step(data, st)
    pos, kin, f = data.positions(st), data.kinematics(st), data.def_grad(st)
    sid_of = data.scene_ids(st)
    assert np.array_equal(sid_of, b.scene_ids)
    for sid in b.ids:
        m = sid_of == sid
        p2, k2, f2 = data.read_scene(sid, st)
        assert np.array_equal(p2[:, :3], pos[m, :3])
        assert np.array_equal(k2, kin[m])
        assert np.array_equal(f2, f[m])
    assert data.scene_size(999) == 0


def test_device_remove_scenes_keeps_others_exact(gpu_pipeline):
    pl, st = gpu_pipeline
    b = small_batch()
    data = P.MpmData.from_batch(pl, b)
    for _ in range(4):
        pl.launch_step(data, st)
    pos, kin, f = data.positions(st), data.kinematics(st), data.def_grad(st)
    models, props = data.models(st), data.properties(st)
    keep = (b.scene_ids != 7) & (b.scene_ids != 12)
    removed = data.remove_scenes([7, 12], st)
    assert removed == int((~keep).sum())
    assert data.n == int(keep.sum())
    assert np.array_equal(data.scene_ids(st), b.scene_ids[keep])
    assert np.array_equal(data.positions(st)[:, :3], pos[keep, :3])
    assert np.array_equal(data.kinematics(st), kin[keep])
    assert np.array_equal(data.def_grad(st), f[keep])
    assert np.array_equal(data.models(st), models[keep])
    assert np.array_equal(data.properties(st), props[keep])
    # The remaining scenes continue as if alone: compare with a batch of only them.
    ref = P.MpmData.from_batch(pl, scenes.Batch([b.scenes[0], b.scenes[2]], ids=[0, 3]))
    for _ in range(4):
        pl.launch_step(ref, st)
    for _ in range(3):
        pl.launch_step(data, st)
        pl.launch_step(ref, st)
    data.check_errors(st)
    # Equal up to summation order within cells (7 substeps).
    for sid in (0, 3):
        a, r = data.read_scene(sid, st), ref.read_scene(sid, st)
        errs = state_errors(a, r, r[0])
        assert max(errs[k] for k in ("x", "v", "affine", "F")) <= tol_for(7), errs


def test_shift_remove_on_device_then_step(gpu_pipeline):
    pl, st = gpu_pipeline
    sc = scenes.elastic_cube(n_side=12)
    data = P.MpmData.from_scene(pl, sc)
    for _ in range(2):
        pl.launch_step(data, st)
    pos = data.positions(st)
    data.shift_remove(100, 900)
    assert data.n == sc.n - 800
    assert np.array_equal(data.positions(st)[:, :3], np.concatenate([pos[:100, :3], pos[900:, :3]]))
    pl.launch_steps_graph(data, 4, st)
    data.check_errors(st)


def test_append_scene_to_running_batch(gpu_pipeline):
    pl, st = gpu_pipeline
    b = small_batch()
    data = P.MpmData.from_batch(pl, b, particle_capacity=b.n + 4000)
    for _ in range(2):
        pl.launch_step(data, st)
    extra = scenes.stir(scenes.batched_scene(9, n_override=3500), 91)
    data.append(extra.particles, scene_ids=np.full(extra.n, 40, np.uint32), stream=st)
    for _ in range(5):
        pl.launch_step(data, st)
    data.check_errors(st)
    orc = O.Oracle(extra, svd_mode=1)
    orc.step(5)
    errs = state_errors(data.read_scene(40, st), (orc.positions(), orc.kinematics(), orc.def_grad()),
                        extra.particles["pos4"])
    assert max(errs[k] for k in ("x", "v", "affine", "F")) <= tol_for(5), errs


def test_autotune_keeps_state_and_results(gpu_pipeline):
    """The autotuner times candidates on a private copy: the data does not advance,
    and a substep under the chosen launch shape matches the default shape."""
    pl, st = gpu_pipeline
    b = small_batch()
    a = P.MpmData.from_batch(pl, b)
    ref = P.MpmData.from_batch(pl, b)
    pl.launch_step(a, st)
    pl.launch_step(ref, st)
    before = a.positions(st)
    t = a.autotune(pl, st, trial_steps=4)
    assert t == a.tuning and t["p2g_chunk"] in (192, 256, 384, 512) and t["g2p_threads"] in (128, 256)
    assert np.array_equal(a.positions(st), before)
    steps = 1
    for shape in ({"p2g_chunk": 256, "g2p_threads": 128, "block_wgs": 2048},
                  {"p2g_chunk": 384, "g2p_threads": 256, "block_wgs": 16384}):
        a.set_tuning(**shape)
        for _ in range(3):
            pl.launch_step(a, st)
            pl.launch_step(ref, st)
        steps += 3
        a.check_errors(st)
        # Equal up to summation order within cells (particle order within a cell
        # follows atomic ranks, which vary between runs).
        for sid in b.ids:
            errs = state_errors(a.read_scene(sid, st), ref.read_scene(sid, st), ref.read_scene(sid, st)[0])
            assert max(errs[k] for k in ("x", "v", "affine", "F")) <= tol_for(steps), (shape, sid, errs)
    with pytest.raises(Exception):
        a.set_tuning(p2g_chunk=300)


def test_batched_key_range_is_reported(gpu_pipeline):
    """A scene whose particles fly out of the batched key range ([-16, 46] blocks per
    axis) is reported by check_errors, and an append outside the range is refused;
    the other scene of the batch is unaffected by the refused append."""
    from slosh_amd import _lib
    from slosh_amd import models as M
    pl, st = gpu_pipeline
    h = 1.0 / 128
    g = np.arange(6) * (h / 2)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    model = M.elastic_neo_hookean(1e5, 0.3)

    def scene(origin, vx):
        parts = M.particle_arrays((lat + origin).astype(np.float32), np.float32(h / 4), 1000.0, model,
                                  np.tile(np.float32([vx, 0.0, 0.0]), (len(lat), 1)))
        return scenes.Scene("range", parts, 1e-4, h, 1024, [], [], gravity=(0.0, 0.0, 0.0))

    # block 45 at the start (x ~ 1.40-1.42), moving at 50 m/s = 0.64 cells per substep:
    # block 47 (x >= 1.4727) after ~10 substeps.
    b = scenes.Batch([scene((1.40, 0.5, 0.5), 50.0), scene((0.3, 0.3, 0.3), 0.0)], grid_capacity=1024)
    data = P.MpmData.from_batch(pl, b)
    for _ in range(3):
        pl.launch_step(data, st)
    data.check_errors(st)  # still in range
    bad = scene((1.48, 0.5, 0.5), 0.0)
    with pytest.raises(_lib.SloshError) as e:
        data.append(bad.particles, np.full(bad.n, 5, np.uint32), st)
    assert e.value.code == -8
    assert data.n == b.n
    for _ in range(20):
        pl.launch_step(data, st)
    with pytest.raises(_lib.SloshError) as e:
        data.check_errors(st)
    assert e.value.code == -8
    data.check_errors(st)  # the flag was cleared
    # the resting scene is untouched by the other's escape
    pos = data.read_scene(1, st, positions_only=True)
    assert np.allclose(pos[:, :3], b.scenes[1].particles["pos4"][:, :3], atol=1e-6)


def _mixed_record_scene():
    """A batched-configuration scene whose particles do not share one static
    record (every other particle gets another Young's modulus)."""
    from slosh_amd import models as M
    sc = scenes.stir(scenes.batched_scene(5, n_override=2500), 93)
    sc.particles["models"][::2] = np.asarray(M.elastic_neo_hookean(5.0e4, 0.3), np.uint32)
    return sc


@pytest.mark.parametrize("steps", [1, 20])
def test_batch_with_a_mixed_record_scene(gpu_pipeline, steps):
    """Scenes of one record each read it from the per-scene table; a scene whose
    particles differ makes every scene read its records by particle id. Both
    against the oracle on each scene alone."""
    pl, st = gpu_pipeline
    scs = [scenes.stir(scenes.batched_scene(0, n_override=2000), 81), _mixed_record_scene(),
           scenes.stir(scenes.batched_scene(2, n_override=3000), 83)]
    b = scenes.Batch(scs, ids=[4, 9, 2])
    data = P.MpmData.from_batch(pl, b)
    for _ in range(steps):
        pl.launch_step(data, st)
    data.check_errors(st)
    for sc, sid in zip(b.scenes, b.ids):
        orc = O.Oracle(sc, svd_mode=1)
        orc.step(steps)
        errs = state_errors(data.read_scene(sid, st), (orc.positions(), orc.kinematics(), orc.def_grad()),
                            sc.particles["pos4"])
        print(f"\n[mixed-record batch, scene {sid} after {steps}] " +
              " ".join(f"{k}={v:.3e}" for k, v in errs.items() if k != "enabled_diff"))
        assert errs["enabled_diff"] == 0
        for k in ("x", "v", "affine", "F"):
            assert errs[k] <= tol_for(steps), (sid, k, errs[k])


def test_scene_id_reused_with_another_material(gpu_pipeline):
    """A removed scene's id taken by a scene of another material (its per-scene
    record is rewritten), and a mixed-record scene appended and removed again
    (by-id records while it runs, the table afterwards): every scene still
    matches the oracle alone."""
    pl, st = gpu_pipeline
    s0 = scenes.stir(scenes.batched_scene(0, n_override=2000), 61)
    s1 = scenes.stir(scenes.batched_scene(1, n_override=1500), 62)
    b = scenes.Batch([s0, s1], ids=[0, 1])
    data = P.MpmData.from_batch(pl, b, particle_capacity=b.n + 8000)
    for _ in range(3):
        pl.launch_step(data, st)
    data.remove_scenes([1], st)
    # Scene id 1 again, with the first scene of another material in the batch configuration.
    tag1 = int(s1.particles["models"][0][0])
    k = next(i for i in range(2, 40) if int(scenes.batched_scene(i, n_override=8).particles["models"][0][0]) != tag1)
    s2 = scenes.stir(scenes.batched_scene(k, n_override=2500), 63)
    data.append(s2.particles, scene_ids=np.full(s2.n, 1, np.uint32), stream=st)
    mx = _mixed_record_scene()
    data.append(mx.particles, scene_ids=np.full(mx.n, 6, np.uint32), stream=st)
    for _ in range(4):
        pl.launch_step(data, st)
    data.remove_scenes([6], st)
    for _ in range(4):
        pl.launch_step(data, st)
    data.check_errors(st)
    for sc, sid, steps in ((s0, 0, 11), (s2, 1, 8)):
        orc = O.Oracle(sc, svd_mode=1)
        orc.step(steps)
        errs = state_errors(data.read_scene(sid, st), (orc.positions(), orc.kinematics(), orc.def_grad()),
                            sc.particles["pos4"])
        print(f"\n[scene {sid} after {steps}] " + " ".join(f"{k}={v:.3e}" for k, v in errs.items()))
        assert errs["enabled_diff"] == 0
        for key in ("x", "v", "affine", "F"):
            assert errs[key] <= tol_for(steps), (sid, key, errs[key])


def _density_scene(i, n, f):
    sc = scenes.stir(scenes.batched_scene(i + 3, n_override=n), 71 + i)
    kin = sc.particles["kin"]
    kin["mass"] = (kin["mass"] * np.float32(f)).astype(np.float32)
    if i == 1:
        kin["enabled"][::97] = 0
        kin["mass"][::97] = 0.0
    return sc


def _oracle_noise_floor(sc, steps, ref):
    """How far the oracle moves when every input position moves by one ulp (the
    HIP and oracle P2G sums differ in summation order, i.e. at that level)."""
    p = sc.particles["pos4"].copy()
    p[:, :3] = np.nextafter(p[:, :3], np.float32(2.0))
    sc2 = copy.copy(sc)
    sc2.particles = dict(sc.particles, pos4=p)
    o = O.Oracle(sc2, svd_mode=1)
    o.step(steps)
    return state_errors((o.positions(), o.kinematics(), o.def_grad()),
                        (ref.positions(), ref.kinematics(), ref.def_grad()), sc.particles["pos4"])


@pytest.mark.parametrize("steps", [1, 20])
def test_batch_scenes_of_different_densities(gpu_pipeline, steps):
    """Scenes whose particles share one mass per scene (but not across scenes):
    the update takes each scene's mass from the per-scene table, and a disabled
    particle (mass +0) keeps a zero mass. Each scene against the oracle alone,
    at the parity gates or, where the oracle itself moves by more than the gate
    under a one-ulp change of its input positions (the half-density scene's APIC
    matrix after one substep: the stress term then dominates m C, 1.3e-5), within
    twice that floor."""
    pl, st = gpu_pipeline
    scs = [_density_scene(i, n, f) for i, (n, f) in enumerate(((2000, 0.5), (1800, 2.0), (2600, 1.0)))]
    b = scenes.Batch(scs, ids=[1, 2, 3])
    data = P.MpmData.from_batch(pl, b)
    for _ in range(steps):
        pl.launch_step(data, st)
    data.check_errors(st)
    for sc, sid in zip(b.scenes, b.ids):
        orc = O.Oracle(sc, svd_mode=1)
        orc.step(steps)
        pos, kin, f = data.read_scene(sid, st)
        assert np.array_equal(kin["mass"], orc.kinematics()["mass"])
        errs = state_errors((pos, kin, f), (orc.positions(), orc.kinematics(), orc.def_grad()),
                            sc.particles["pos4"])
        floor = _oracle_noise_floor(sc, steps, orc)
        print(f"\n[density batch, scene {sid} after {steps}] " +
              " ".join(f"{k}={v:.3e}" for k, v in errs.items() if k != "enabled_diff") +
              " | oracle floor " + " ".join(f"{k}={floor[k]:.2e}" for k in ("x", "v", "affine", "F")))
        assert errs["enabled_diff"] == 0
        for k in ("x", "v", "affine", "F"):
            assert errs[k] <= max(tol_for(steps), 2.0 * floor[k]), (sid, k, errs[k], floor[k])


def test_appends_staged_through_a_small_ring(gpu_pipeline):
    """Appends while substeps are queued, staged in a 600 KB pinned ring (200 bytes
    per particle: regions wrap around it and wait for the oldest uploads), one
    append too large for the ring (staged in an arena instead), and one scene with
    disabled particles: every scene matches the oracle run alone for its substeps."""
    pl, st = gpu_pipeline
    first = scenes.stir(scenes.batched_scene(0, n_override=1500), 51)
    data = P.MpmData.from_batch(pl, scenes.Batch([first], ids=[0]), particle_capacity=30000)
    data.reserve_staging(600_000)
    runs = [(first, 0, 0)]  # scene, id, substep at which it joined
    sizes = (1200, 2400, 1800, 4000, 2200, 1000)
    step = 0
    for k, n in enumerate(sizes):
        for _ in range(2):
            pl.launch_step(data, st)
            step += 1
        sc = scenes.stir(scenes.batched_scene(k + 1, n_override=n), 52 + k)
        if k == 2:
            kin = sc.particles["kin"]
            kin["enabled"][::13] = 0
            kin["mass"][::13] = 0.0
        data.append(sc.particles, scene_ids=np.full(sc.n, k + 1, np.uint32), stream=st)
        runs.append((sc, k + 1, step))
    for _ in range(3):
        pl.launch_step(data, st)
        step += 1
    data.check_errors(st)
    for sc, sid, joined in runs:
        steps = step - joined
        orc = O.Oracle(sc, svd_mode=1)
        orc.step(steps)
        pos, kin, f = data.read_scene(sid, st)
        assert np.array_equal(kin["mass"], orc.kinematics()["mass"])
        errs = state_errors((pos, kin, f), (orc.positions(), orc.kinematics(), orc.def_grad()),
                            sc.particles["pos4"])
        print(f"\n[ring append, scene {sid} n={sc.n} after {steps}] " +
              " ".join(f"{k}={v:.3e}" for k, v in errs.items() if k != "enabled_diff"))
        assert errs["enabled_diff"] == 0
        for key in ("x", "v", "affine", "F"):
            assert errs[key] <= tol_for(steps), (sid, key, errs[key])
