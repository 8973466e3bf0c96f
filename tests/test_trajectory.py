"""Trajectory format (validation_tests/src/harness.rs:25-109, validation_tests/README.md
"JSON Trajectory Format"): JSON round trip, CSV layout, and the README's example document."""
import json
import os

import numpy as np
import pytest

from slosh_amd import trajectory as T

README_EXAMPLE = """{
  "name": "elastic_beam_cantilever",
  "dt": 0.0167,
  "num_substeps": 20,
  "snapshots": [
    {
      "time": 0.0,
      "step": 0,
      "particles": [
        {"position": [1.0, 2.0, 3.0], "velocity": [0.0, 0.0, 0.0]}
      ]
    }
  ],
  "metadata": {
    "num_particles": 1000,
    "cell_width": 0.5,
    "gravity": [0.0, -9.81, 0.0],
    "material": {
      "young_modulus": 1000000.0,
      "poisson_ratio": 0.3,
      "density": 1000.0,
      "material_type": "neo_hookean"
    }
  }
}"""


def sample(n=50, steps=(0, 1, 3), seed=5):
    rng = np.random.default_rng(seed)
    snaps = []
    for k, st in enumerate(steps):
        pos = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(10.0 ** rng.integers(-6, 3))
        vel = rng.normal(size=(n, 3)).astype(np.float32)
        snaps.append(T.SimulationSnapshot(float(np.float32(st * 1.0 / 60)), st, pos, vel))
    meta = T.SimulationMetadata(n, float(np.float32(1 / 128)), (0.0, float(np.float32(-9.81)), 0.0),
                                T.MaterialParams(1.0e6, float(np.float32(0.3)), 1000.0, "neo_hookean"))
    return T.SimulationTrajectory("sample", float(np.float32(1 / 60)), 20, snaps, meta)


def test_json_round_trip_is_exact(tmp_path):
    tr = sample()
    path = tmp_path / "t.json"
    tr.export_json(path)
    back = T.SimulationTrajectory.load_json(path)
    assert back.name == tr.name and back.num_substeps == tr.num_substeps
    assert np.float32(back.dt) == np.float32(tr.dt)
    assert len(back.snapshots) == len(tr.snapshots)
    for a, b in zip(tr.snapshots, back.snapshots):
        assert a.step == b.step and np.float32(a.time) == np.float32(b.time)
        assert np.array_equal(a.positions, b.positions)  # bit-exact float32 round trip
        assert np.array_equal(a.velocities, b.velocities)
    assert back.metadata == tr.metadata


def test_json_has_the_reference_structure():
    d = json.loads(sample(n=2, steps=(0,)).to_json())
    assert list(d) == ["name", "dt", "num_substeps", "snapshots", "metadata"]
    assert list(d["snapshots"][0]) == ["time", "step", "particles"]
    assert list(d["snapshots"][0]["particles"][0]) == ["position", "velocity"]
    assert list(d["metadata"]) == ["num_particles", "cell_width", "gravity", "material"]
    assert list(d["metadata"]["material"]) == ["young_modulus", "poisson_ratio", "density", "material_type"]


def test_reads_the_readme_example():
    tr = T.SimulationTrajectory.from_json(README_EXAMPLE)
    assert tr.name == "elastic_beam_cantilever" and tr.num_substeps == 20
    assert np.float32(tr.dt) == np.float32(0.0167)
    assert np.array_equal(tr.snapshots[0].positions, np.array([[1.0, 2.0, 3.0]], np.float32))
    assert tr.metadata.material.material_type == "neo_hookean"
    assert tr.metadata.num_particles == 1000
    # and writes it back to the same values
    again = json.loads(tr.to_json())
    assert again == json.loads(README_EXAMPLE)


def test_csv_layout(tmp_path):
    tr = sample(n=7)
    tr.export_csv(tmp_path / "csv")
    names = sorted(os.listdir(tmp_path / "csv"))
    assert names == ["step_000000.csv", "step_000001.csv", "step_000003.csv"]
    lines = open(tmp_path / "csv" / "step_000003.csv").read().splitlines()
    assert lines[0] == "particle_id,x,y,z,vx,vy,vz"
    assert len(lines) == 8 and lines[1].startswith("0,") and lines[7].startswith("6,")
    ids, pos, vel = T.load_csv(tmp_path / "csv" / "step_000003.csv")
    assert np.array_equal(ids, np.arange(7))
    assert np.array_equal(pos, tr.snapshots[2].positions) and np.array_equal(vel, tr.snapshots[2].velocities)


def test_number_text_is_shortest_float32():
    # Rust prints f32 with the shortest round-trip decimal and no exponent.
    assert T._f32(1.0) == "1" and T._f32(0.5) == "0.5"
    assert T._f32(np.float32(0.1)) == "0.1"
    assert T._f32(1e-5) == "0.00001"
    with pytest.raises(ValueError):
        T._f32(float("nan"))
