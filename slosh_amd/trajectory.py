"""Simulation trajectories in the reference's validation format.

Mirrors ``SimulationTrajectory`` and its JSON / CSV exports
(validation_tests/src/harness.rs:25-109; the format is documented in
validation_tests/README.md "JSON Trajectory Format"):

* JSON: ``{"name", "dt", "num_substeps", "snapshots": [{"time", "step",
  "particles": [{"position": [x, y, z], "velocity": [vx, vy, vz]}, ...]}, ...],
  "metadata": {"num_particles", "cell_width", "gravity", "material":
  {"young_modulus", "poisson_ratio", "density", "material_type"}}}``
* CSV: one file per snapshot, ``step_{step:06}.csv``, header
  ``particle_id,x,y,z,vx,vy,vz``, one row per particle in particle order.

``record`` runs the loop of ``ValidationHarness::run_scenario``
(harness.rs:182-276): a snapshot of the initial state, then ``total_steps``
steps of ``num_substeps`` substeps each (substep dt = dt / num_substeps), with a
snapshot every ``snapshot_interval`` steps. The reference records zero
velocities ("would need a custom readback shader", harness.rs:307); here the
particles' velocities are read back with their positions (``zero_velocity=True``
writes zeros like the reference). Values are float32; text output uses the
shortest decimal that reads back to the same float32 (the reference writes f32
with Rust's shortest round-trip formatting).
"""
import dataclasses
import json
import os

import numpy as np

from . import layouts as L


@dataclasses.dataclass
class MaterialParams:
    young_modulus: float
    poisson_ratio: float
    density: float
    material_type: str


@dataclasses.dataclass
class SimulationMetadata:
    num_particles: int
    cell_width: float
    gravity: tuple
    material: MaterialParams


@dataclasses.dataclass
class SimulationSnapshot:
    time: float
    step: int
    positions: np.ndarray   # (n, 3) float32
    velocities: np.ndarray  # (n, 3) float32


@dataclasses.dataclass
class SimulationTrajectory:
    name: str
    dt: float
    num_substeps: int
    snapshots: list
    metadata: SimulationMetadata

    # -- JSON ---------------------------------------------------------------
    def to_json(self):
        snaps = []
        for s in self.snapshots:
            parts = ",".join('{"position":[%s],"velocity":[%s]}' % (_f32_list(p), _f32_list(v))
                             for p, v in zip(s.positions, s.velocities))
            snaps.append('{"time":%s,"step":%d,"particles":[%s]}' % (_f32(s.time), int(s.step), parts))
        m = self.metadata
        meta = ('{"num_particles":%d,"cell_width":%s,"gravity":[%s],"material":{"young_modulus":%s,'
                '"poisson_ratio":%s,"density":%s,"material_type":%s}}'
                % (int(m.num_particles), _f32(m.cell_width), _f32_list(m.gravity), _f32(m.material.young_modulus),
                   _f32(m.material.poisson_ratio), _f32(m.material.density), json.dumps(m.material.material_type)))
        return '{"name":%s,"dt":%s,"num_substeps":%d,"snapshots":[%s],"metadata":%s}' % (
            json.dumps(self.name), _f32(self.dt), int(self.num_substeps), ",".join(snaps), meta)

    def export_json(self, path):
        """SimulationTrajectory::export_json (harness.rs:71-76)."""
        with open(path, "w") as f:
            f.write(self.to_json())

    @staticmethod
    def from_json(text):
        d = json.loads(text)
        snaps = []
        for s in d["snapshots"]:
            parts = s["particles"]
            pos = np.array([p["position"] for p in parts], dtype=np.float32).reshape(-1, 3)
            vel = np.array([p["velocity"] for p in parts], dtype=np.float32).reshape(-1, 3)
            snaps.append(SimulationSnapshot(float(np.float32(s["time"])), int(s["step"]), pos, vel))
        m = d["metadata"]
        mat = m["material"]
        meta = SimulationMetadata(int(m["num_particles"]), float(np.float32(m["cell_width"])),
                                  tuple(float(np.float32(g)) for g in m["gravity"]),
                                  MaterialParams(float(np.float32(mat["young_modulus"])),
                                                 float(np.float32(mat["poisson_ratio"])),
                                                 float(np.float32(mat["density"])), str(mat["material_type"])))
        return SimulationTrajectory(str(d["name"]), float(np.float32(d["dt"])), int(d["num_substeps"]), snaps, meta)

    @staticmethod
    def load_json(path):
        """SimulationTrajectory::load_json (harness.rs:103-108)."""
        with open(path) as f:
            return SimulationTrajectory.from_json(f.read())

    # -- CSV ----------------------------------------------------------------
    def export_csv(self, directory):
        """SimulationTrajectory::export_csv (harness.rs:78-101): step_{:06}.csv per snapshot."""
        os.makedirs(directory, exist_ok=True)
        for s in self.snapshots:
            rows = ["particle_id,x,y,z,vx,vy,vz"]
            for i, (p, v) in enumerate(zip(s.positions, s.velocities)):
                rows.append(",".join([str(i)] + [_f32(c) for c in p] + [_f32(c) for c in v]))
            with open(os.path.join(directory, "step_%06d.csv" % s.step), "w") as f:
                f.write("\n".join(rows) + "\n")


def load_csv(path):
    """One CSV snapshot: (particle ids, positions (n, 3), velocities (n, 3)) as float32."""
    with open(path) as f:
        header = f.readline().strip().split(",")
        if header != ["particle_id", "x", "y", "z", "vx", "vy", "vz"]:
            raise ValueError(f"{path}: not a trajectory snapshot (header {header})")
        rows = [line.strip().split(",") for line in f if line.strip()]
    ids = np.array([int(r[0]) for r in rows], dtype=np.int64)
    vals = np.array([[np.float32(c) for c in r[1:]] for r in rows], dtype=np.float32).reshape(-1, 6)
    return ids, vals[:, :3], vals[:, 3:]


def _f32(x):
    """Shortest decimal that reads back to the same float32 (no exponent)."""
    x = np.float32(x)
    if not np.isfinite(x):
        raise ValueError("non-finite value in a trajectory")
    return np.format_float_positional(x, unique=True, trim="-")


def _f32_list(v):
    return ",".join(_f32(c) for c in v)


def snapshot(data, step, time, stream=None, zero_velocity=False):
    """ValidationHarness::extract_snapshot (harness.rs:279-312): positions in particle order."""
    pos = data.positions(stream)[:, :3].astype(np.float32)
    if zero_velocity:
        vel = np.zeros_like(pos)
    else:
        vel = np.ascontiguousarray(data.kinematics(stream)["velocity"], dtype=np.float32).reshape(-1, 3)
    return SimulationSnapshot(float(np.float32(time)), int(step), pos, vel)


def record(pipeline, data, name, dt, num_substeps, total_steps, material, snapshot_interval=1, stream=None,
           zero_velocity=False):
    """ValidationHarness::run_scenario (harness.rs:182-276) on resident data.

    ``dt`` is the step (frame) duration; ``data`` must have been created with
    substep dt = dt / num_substeps (SimulationParams, harness.rs:189-192).
    ``material`` is a MaterialParams for the metadata.
    """
    params = data.params
    sub_dt = np.float32(np.float32(dt) / np.float32(num_substeps))
    if np.float32(params["dt"]) != sub_dt:
        raise ValueError(f"data substep dt {float(params['dt'])} != dt / num_substeps = {float(sub_dt)}")
    snaps = [snapshot(data, 0, 0.0, stream, zero_velocity)]
    t = np.float32(0.0)
    for step in range(1, total_steps + 1):
        pipeline.launch_steps_graph(data, num_substeps, stream)
        t = np.float32(t + np.float32(dt))
        if step % snapshot_interval == 0:
            snaps.append(snapshot(data, step, t, stream, zero_velocity))
    meta = SimulationMetadata(int(data.n), float(np.float32(data.cell_width)),
                              tuple(float(np.float32(g)) for g in params["gravity"]), material)
    return SimulationTrajectory(str(name), float(np.float32(dt)), int(num_substeps), snaps, meta)


__all__ = ["MaterialParams", "SimulationMetadata", "SimulationSnapshot", "SimulationTrajectory", "snapshot",
           "record", "load_csv", "L"]
