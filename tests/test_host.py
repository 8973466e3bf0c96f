"""Host-side logic that needs no GPU: scene batching, the request stream, record
layouts handed to the C ABI."""
import numpy as np

from slosh_amd import layouts as L
from slosh_amd import models as M
from slosh_amd import pipeline as P
from slosh_amd import scenes
from slosh_amd import serving as SV


def test_batch_keeps_record_layouts_and_scene_ids():
    scs = [scenes.batched_scene(i, n_override=n) for i, n in enumerate((100, 250, 40))]
    b = scenes.Batch(scs, ids=[5, 2, 9])
    assert b.n == 390
    assert b.particles["kin"].dtype == L.KINEMATICS
    assert b.particles["props"].dtype == L.PROPERTIES
    assert np.array_equal(b.scene_ids, np.repeat([5, 2, 9], [100, 250, 40]))
    for k in b.particles:
        assert np.array_equal(b.particles[k][100:350], scs[1].particles[k])
    st, keep = P._particles_struct(b.particles, b.scene_ids)
    assert st.len == 390 and st.scene_ids


def test_concat_particles_preserves_padded_dtype():
    a = M.particle_arrays(np.zeros((3, 3), np.float32), np.float32(0.01), 1000.0, M.elastic(1e5, 0.3))
    c = M.concat_particles([a, a])
    assert c["kin"].dtype.itemsize == 96 and c["kin"].dtype == a["kin"].dtype
    assert c["pos4"].shape == (6, 4)


def test_empty_batched_particles_mark_batched():
    st, keep = P._particles_struct(SV._empty_particles(), np.zeros(0, np.uint32))
    assert st.len == 0 and st.scene_ids  # non-null: batched even when empty


def test_batch_rejects_mismatched_colliders():
    a = scenes.batched_scene(0, n_override=50)
    b = scenes.elastic_cube(n_side=4)
    try:
        scenes.Batch([a, b])
    except AssertionError:
        return
    raise AssertionError("a batch with different cell widths was accepted")


def test_block_estimate_bounds_a_lattice():
    sc = scenes.batched_scene(3, n_override=4000)
    pos = sc.particles["pos4"][:, :3]
    h = float(sc.cell_width)
    # Brute force: the base blocks of every particle plus their +1 neighbours.
    base = np.floor((np.rint(pos / h) - 1) / 4).astype(int)
    blocks = {tuple(b + d) for b in np.unique(base, axis=0)
              for d in np.array(np.meshgrid([0, 1], [0, 1], [0, 1])).T.reshape(-1, 3)}
    assert len(blocks) <= scenes.estimate_blocks(sc.particles["pos4"], sc.cell_width)


def test_request_stream_is_seeded_and_in_range():
    a = [scenes.served_request(k) for k in range(20)]
    b = [scenes.served_request(k) for k in range(20)]
    for (sa, ka), (sb, kb) in zip(a, b):
        assert ka == kb and np.array_equal(sa.particles["pos4"], sb.particles["pos4"])
        assert 5000 <= sa.n <= 2_000_000 and 100 <= ka <= 2000
    t = scenes.poisson_arrivals(50.0, 4000)
    assert np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 1 / 50.0) < 0.1 / 50.0


def test_served_latency_summary():
    class R:
        def __init__(self, n, s, a, f):
            self.n, self.substeps, self.arrival, self.finished = n, s, a, f
            self.latency = f - a
    rs = [R(100, 10, 0.0, 1.0), R(200, 10, 0.5, 2.0), R(50, 20, 1.0, 1.5)]
    s = SV.latency_summary(rs)
    assert s["requests"] == 3 and s["particle_substeps"] == 4000
    assert abs(s["makespan_s"] - 2.0) < 1e-12 and abs(s["p50_latency_s"] - 1.0) < 1e-12


def test_bench_refuses_other_gpu_counts_and_empty_runs():
    """bench.py's command line: one GPU only (a JSON line saying so, exit code 2,
    before anything touches a GPU) and at least one timed substep."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--steps", "0"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "--steps must be at least 1" in r.stderr
