"""Served scenes: a continuous-batching request loop over one device-resident batch.

Scene requests (a scene plus a substep count) arrive over time. Every admitted
scene lives in one batched ``MpmData`` under its own scene id (its own block-key
space), so each substep of the batch advances every running scene in the same
kernel launches. Between chunks of substeps the loop

* retires the scenes that reached their substep count: their particles are
  gathered on the device (``slosh_read_scene``), copied out as the result, and
  removed by an on-device compaction (``slosh_remove_scenes``);
* admits queued requests first-come first-served while the batch's particle
  capacity, its hash-grid capacity and the HBM budget allow, by uploading only the
  new particles into the free tail of the resident storage
  (``slosh_append_particles``).

A chunk runs as many substeps as the running scene closest to completion still
needs (at most ``max_chunk``), so every scene stops at exactly its own count.
Scenes never interact (disjoint key spaces; shared colliders are fixed), so a
scene's result is the one it gets alone, up to the summation order of the
particle-to-grid sums within a cell.

The reference has no serving layer; this is the request loop around its
``MpmPipeline::launch_step`` (src/pipeline.rs:562-770) with ``MpmData`` kept
resident across requests.
"""
import collections
import time

import numpy as np

from . import layouts as L
from . import pipeline as P
from . import scenes as S

MAX_SCENES = 16383

# Device bytes per particle slot of a batched MpmData: two particle buffers
# (pos 16, vel+mass 16, C 36, F 36, vel_grad_det 4, id 4, Drucker-Prager state 12),
# force_dt 12 x 2, by-id records 64 x 2, full-sort scratch 12 (slot, rank,
# permutation), mover lists 16 (key, slot, overflow slot and block).
BYTES_PER_PARTICLE = 2 * (16 + 16 + 36 + 36 + 4 + 4 + 12) + 2 * 12 + 2 * 64 + 12 + 16
# Device bytes per hash-grid slot: per table (two) the key, block start, count,
# stayers and movers (20 x 2), the block-list entry 4, neighbour table 64, cell
# counts and starts 512, mover list 64 x 4, previous range 16, P2G tile 216 x 16,
# nodes 64 x 16.
BYTES_PER_BLOCK = 2 * 20 + 4 + 64 + 512 + 256 + 16 + 216 * 16 + 64 * 16
# Pinned host bytes an append stages one particle in.
STAGING_BYTES_PER_PARTICLE = 200


class Request:
    """One scene request: the scene, its substep count, its arrival time (s)."""

    def __init__(self, rid, scene, substeps, arrival=0.0):
        self.rid = rid
        self.scene = scene
        self.substeps = int(substeps)
        self.arrival = float(arrival)
        self.blocks = S.estimate_blocks(scene.particles["pos4"], scene.cell_width)
        self.admitted = None
        self.finished = None
        self.result = None

    @property
    def n(self):
        return self.scene.n

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def _empty_particles():
    return {"pos4": np.zeros((0, 4), np.float32), "kin": np.zeros(0, L.KINEMATICS),
            "def_grad": np.zeros((0, 3, 4), np.float32), "props": np.zeros(0, L.PROPERTIES),
            "models": np.zeros((0, 13), np.uint32)}


class SceneServer:
    """Continuous batching of scene requests on one device-resident batch.

    template        a Scene whose time step, gravity, cell width and colliders every
                    request shares (one launch applies them to all scenes)
    hbm_budget      device bytes the batch may hold (particle and grid storage)
    max_particles   particles in flight (bounds the substep time, hence latency)
    grid_capacity   hash-grid slots (power of two); admission keeps the blocks
                    reserved for running scenes under grid_load of it
    block_growth    blocks reserved per scene = block_growth x its initial estimate
                    (scenes spread while they run)
    result          "positions" (default) or "full" (positions, kinematics, F)
    staging_bytes   pinned host ring reserved at construction for admissions to stage
                    in (default: 200 bytes per particle of up to 16M particles), so
                    that admitting a request pins no memory
    use_graphs      replay captured substep pairs instead of launching each substep.
                    Off by default: the pairs are re-captured after every change of
                    the batch's membership, and a batch of millions of particles
                    takes far longer per substep than its launches (measured 3-4 %
                    higher throughput without graphs on the served stream)
    """

    def __init__(self, pipeline, template, hbm_budget=200e9, max_particles=24 << 20, grid_capacity=1 << 20,
                 grid_load=0.7, block_growth=3.0, max_chunk=64, result="positions", stream=None, use_graphs=False,
                 staging_bytes=None):
        grid_capacity = S.pow2_at_least(grid_capacity)
        grid_bytes = grid_capacity * BYTES_PER_BLOCK
        if grid_bytes >= hbm_budget:
            raise ValueError("HBM budget smaller than the hash grid")
        self.particle_capacity = int(min(max_particles, (hbm_budget - grid_bytes) // BYTES_PER_PARTICLE))
        self.grid_capacity = grid_capacity
        self.block_budget = int(grid_load * grid_capacity)
        self.block_growth = block_growth
        self.device_bytes = self.particle_capacity * BYTES_PER_PARTICLE + grid_bytes
        self.max_chunk = max_chunk
        self.use_graphs = use_graphs
        self.result_kind = result
        self.pipeline = pipeline
        self.stream = stream or P.Stream()
        self.data = P.MpmData(pipeline, template.params, _empty_particles(), template.bodies, template.materials,
                              template.cell_width, grid_capacity, scene_ids=np.zeros(0, np.uint32),
                              particle_capacity=self.particle_capacity)
        if staging_bytes is None:
            staging_bytes = STAGING_BYTES_PER_PARTICLE * min(self.particle_capacity, 16 << 20)
        if staging_bytes:
            self.data.reserve_staging(staging_bytes)
        self.cell_width = template.cell_width
        self.free_ids = collections.deque(range(MAX_SCENES))
        self.queue = collections.deque()
        self.running = {}  # scene id -> [request, substeps left]
        self.reserved_blocks = 0
        self.in_flight = 0
        self.peak_in_flight = 0
        self.substeps_run = 0
        self.particle_substeps = 0
        self.chunks = 0
        self.seconds = {"admit": 0.0, "run": 0.0, "retire": 0.0, "idle": 0.0}
        # Measured GPU seconds per particle-substep of a chunk (0 until the first chunk).
        self.sec_per_ps = 0.0

    # -- admission / retirement -------------------------------------------------
    def submit(self, req):
        if req.n > self.particle_capacity:
            raise ValueError(f"request {req.rid}: {req.n} particles exceed the batch capacity")
        self.queue.append(req)

    def _reservation(self, req):
        return int(np.ceil(self.block_growth * req.blocks))

    def admit(self, now, budget=None):
        """Admits queued requests (FIFO) while particles, grid slots and scene ids
        allow, and, with a time budget (seconds), only while the host time spent
        admitting stays within it (at least one request per call)."""
        t0 = time.perf_counter()
        admitted = []
        while self.queue and self.free_ids:
            if budget is not None and admitted and time.perf_counter() - t0 >= budget:
                break
            req = self.queue[0]
            blocks = self._reservation(req)
            if self.in_flight + req.n > self.particle_capacity:
                break
            if self.reserved_blocks + blocks > self.block_budget and self.running:
                break
            self.queue.popleft()
            sid = self.free_ids.popleft()
            self.data.append(req.scene.particles, scene_ids=np.full(req.n, sid, np.uint32), stream=self.stream)
            req.admitted = now
            req.scene_id = sid
            self.running[sid] = [req, req.substeps]
            self.reserved_blocks += blocks
            self.in_flight += req.n
            admitted.append(req)
        self.peak_in_flight = max(self.peak_in_flight, self.in_flight)
        self.seconds["admit"] += time.perf_counter() - t0
        return admitted

    def retire(self, clock):
        """Reads back and removes every scene whose substeps are done."""
        done = [sid for sid, (_, left) in self.running.items() if left == 0]
        if not done:
            return []
        t0 = time.perf_counter()
        out = []
        for sid in done:
            req = self.running[sid][0]
            if self.result_kind == "full":
                req.result = self.data.read_scene(sid, self.stream)
            else:
                req.result = self.data.read_scene(sid, self.stream, positions_only=True)[:, :3]
        self.data.remove_scenes(done, self.stream)
        now = clock()
        for sid in done:
            req = self.running.pop(sid)[0]
            req.finished = now
            self.free_ids.append(sid)
            self.reserved_blocks -= self._reservation(req)
            self.in_flight -= req.n
            out.append(req)
        self.seconds["retire"] += time.perf_counter() - t0
        return out

    # -- stepping ---------------------------------------------------------------
    def launch_chunk(self, until=None):
        """Enqueues min(substeps left of any running scene, max_chunk) substeps of the
        batch (no further than substep count `until` if given) without waiting;
        returns the token finish_chunk takes."""
        t0 = time.perf_counter()
        self._chunk_t0 = t0
        k = min(min(left for _, left in self.running.values()), self.max_chunk)
        if until is not None and until > self.substeps_run:
            k = min(k, int(until) - self.substeps_run)
        if self.use_graphs:
            self.pipeline.launch_steps_graph(self.data, k, self.stream)
        else:
            for _ in range(k):
                self.pipeline.launch_step(self.data, self.stream)
        self.seconds["run"] += time.perf_counter() - t0
        return k, list(self.running.values()), self.in_flight

    def finish_chunk(self, token):
        """Waits for a launched chunk and advances the scenes it ran."""
        t0 = time.perf_counter()
        k, entries, in_flight = token
        self.stream.synchronize()
        if k * in_flight > 0:
            self.sec_per_ps = (time.perf_counter() - self._chunk_t0) / (k * in_flight)
        self.data.check_errors(self.stream)
        for entry in entries:
            entry[1] -= k
        self.substeps_run += k
        self.particle_substeps += k * in_flight
        self.chunks += 1
        self.seconds["run"] += time.perf_counter() - t0
        return k

    def run_chunk(self, until=None):
        return self.finish_chunk(self.launch_chunk(until))

    def serve(self, requests, realtime=True):
        """Serves `requests`. With realtime, a request becomes visible at its arrival
        time (seconds on the wall clock from the call) and times are seconds;
        otherwise the clock counts the batch's substeps, a request becomes visible
        when that count reaches its arrival, and latencies are in substeps.
        Returns the finished requests."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float(self.substeps_run))
        finished = []
        while pending or self.queue or self.running:
            now = clock()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            # On the wall clock, queued requests are admitted while a chunk runs
            # (below); here only when nothing runs, and then only for a short
            # while, so that the GPU starts on the first of a burst at once.
            if not realtime or not self.running:
                self.admit(now, budget=0.005 if realtime else None)
            if not self.running:
                if pending and realtime:
                    t1 = time.perf_counter()
                    time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.05)))
                    self.seconds["idle"] += time.perf_counter() - t1
                elif pending:
                    self.substeps_run = max(self.substeps_run, int(np.ceil(pending[0].arrival)))
                elif self.queue:
                    raise RuntimeError("queued request can never be admitted")
                continue
            # Without a wall clock, stop at the next arrival so it joins on time.
            token = self.launch_chunk(None if realtime or not pending else np.ceil(pending[0].arrival))
            if realtime:
                # While the GPU runs the chunk, stage the requests that arrived
                # meanwhile (their particles go to free slots the chunk does not
                # touch; they join from the next chunk), for about as long as the
                # chunk runs: a burst is admitted over several chunks, each larger
                # than the last, instead of while the GPU waits.
                now = clock()
                while pending and pending[0].arrival <= now:
                    self.submit(pending.popleft())
                k, _, in_flight = token
                budget = 0.9 * self.sec_per_ps * k * in_flight if self.sec_per_ps > 0.0 else 0.005
                self.admit(now, budget=budget)
            self.finish_chunk(token)
            finished += self.retire(clock)
        return finished


def serve_one_at_a_time(pipeline, requests, template=None, stream=None, realtime=True):
    """The plain per-scene loop the batched server is compared with: requests in
    arrival order, each stepped alone in its own MpmData (captured graphs), read
    back and freed before the next starts."""
    st = stream or P.Stream()
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    for req in pending:
        now = time.perf_counter() - t0
        if realtime and req.arrival > now:
            time.sleep(req.arrival - now)
        req.admitted = time.perf_counter() - t0
        sc = req.scene
        cap = S.pow2_at_least(int(np.ceil(3.0 * req.blocks)) + 256)
        data = P.MpmData(pipeline, sc.params, sc.particles, sc.bodies, sc.materials, sc.cell_width, cap)
        pipeline.launch_steps_graph(data, req.substeps, st)
        st.synchronize()
        data.check_errors(st)
        req.result = data.positions(st)[:, :3].copy()
        data.close()
        req.finished = time.perf_counter() - t0
    return pending


def latency_summary(finished):
    lat = np.array([r.latency for r in finished])
    work = sum(r.n * r.substeps for r in finished)
    first = min(r.arrival for r in finished)
    last = max(r.finished for r in finished)
    return {
        "requests": len(finished),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "makespan_s": float(last - first),
        "particle_substeps": int(work),
        "throughput": float(work / (last - first)),
    }
