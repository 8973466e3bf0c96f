"""Trajectory recording on the GPU (ValidationHarness::run_scenario loop,
validation_tests/src/harness.rs:182-276) through the C ABI."""
import numpy as np
import pytest

from slosh_amd import pipeline as P
from slosh_amd import scenes
from slosh_amd import trajectory as T
from tests.parity import rel_l2, tol_for

pytestmark = pytest.mark.gpu


def test_record_matches_stepping_and_round_trips(gpu_pipeline, tmp_path):
    pl, st = gpu_pipeline
    sc = scenes.stir(scenes.elastic_cube(n_side=10), 31)
    sub_dt = np.float32(sc.params["dt"])
    data = P.MpmData.from_scene(pl, sc)
    mat = T.MaterialParams(1.0e5, 0.3, 1000.0, "neo_hookean")
    tr = T.record(pl, data, "cube", float(sub_dt * np.float32(2)), 2, 4, mat, snapshot_interval=2, stream=st)
    assert [s.step for s in tr.snapshots] == [0, 2, 4]
    assert np.array_equal(tr.snapshots[0].positions, sc.particles["pos4"][:, :3])
    assert tr.metadata.num_particles == sc.n
    # The same substeps run one by one give the same snapshots (up to summation order).
    ref = P.MpmData.from_scene(pl, sc)
    done = 0
    for snap in tr.snapshots[1:]:
        while done < 2 * snap.step:
            pl.launch_step(ref, st)
            done += 1
        assert rel_l2(snap.positions, ref.positions(st)[:, :3]) <= tol_for(done)
        assert rel_l2(snap.velocities, ref.kinematics(st)["velocity"]) <= tol_for(done)
    path = tmp_path / "cube.json"
    tr.export_json(path)
    back = T.SimulationTrajectory.load_json(path)
    for a, b in zip(tr.snapshots, back.snapshots):
        assert np.array_equal(a.positions, b.positions) and np.array_equal(a.velocities, b.velocities)
    with pytest.raises(ValueError):
        T.record(pl, data, "cube", float(sub_dt), 2, 1, mat, stream=st)  # dt / num_substeps mismatch
