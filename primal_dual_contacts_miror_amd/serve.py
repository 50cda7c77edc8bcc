"""Served scene stepping: continuous batching over device-resident scenes.

No counterpart in the reference, which runs one scene per process.

Scenes become resident (their bodies uploaded once, their state kept in HBM)
and leave again; a request asks for one timestep of one resident scene.

  shards     the resident scenes are spread over `shards` contexts, each with
             its own stream and its own worker thread; rounds of different
             shards run on the device at the same time.
  rounds     a shard's worker takes every pending request of its scenes (at
             most one per scene; a second request for a scene waits for the
             next round) and steps those scenes together in one call
             (pdc_step_scenes).  Requests that arrive while a round runs are
             gathered by the calling thread and form the next round, so the
             host never holds the device back.
  segments   a scene occupies a segment of its shard's body table: its bodies
             and one inert padding body (static, far away, emitting no
             samples) behind them.  When a scene leaves, its segment is free,
             and a later scene that fits (bodies and surface samples) is
             written into it in place (pdc_replace_bodies); otherwise the scene
             is appended as a new segment (pdc_append_bodies).  Padding bodies
             fill the rest of a reused segment.  Inert bodies change no result,
             and a scene's work shapes follow its own dynamic or emitting
             bodies, so a scene's result is the same alone and in any mix.
  admission  by HBM: each resident scene holds its footprint (its grids,
             samples and body rows and its expected contacts, scene_footprint)
             in its shard's share of the budget, until it leaves.  A scene is
             admitted to the shard with the most room when its footprint fits
             and the resident count is below `max_resident`; a shard whose
             device bytes (pdc_device_bytes, as they stand) have reached its
             share takes scenes only into free segments.  Otherwise the scene
             waits, and waiting scenes are admitted in arrival order as
             departures free room.
"""
import collections
import threading
import time

import numpy as np

from . import pdc
from .params import PCG_FP64, SHAPE_ELLIPSOID, SHAPE_MESH, SHAPE_SPHERE

PAD_HALF = 1e-3      # half extent of a padding body
PAD_FAR = 1.0e6      # its distance from everything else
CONTACTS_PER_BODY = 16.0
BYTES_PER_CONTACT = 720.0


def sample_count(shape, dims, spacing):
    """Surface samples of one body, by the library's rule (boxes: a cell-centred
    grid of max(1, round(edge / spacing)) per edge on each face; spheres: a
    Fibonacci lattice of max(8, round(4 pi r^2 / spacing^2)) points;
    ellipsoids: that lattice with 4 pi (ab + ac + bc) / 3 in place of the
    sphere's area)."""
    def rnd(x):  # half away from zero, as the library rounds
        return int(np.floor(x + 0.5))
    if shape == SHAPE_SPHERE:
        return max(8, rnd(4.0 * np.pi * dims[0] ** 2 / spacing ** 2))
    if shape == SHAPE_ELLIPSOID:
        a, b, c = dims
        return max(8, rnd(4.0 * np.pi * ((a * b + a * c + b * c) / 3.0) / spacing ** 2))
    e = [max(1, rnd(2.0 * d / spacing)) for d in dims]
    return 2 * (e[1] * e[2] + e[2] * e[0] + e[0] * e[1])


def scene_samples(scene, params=None):
    """Surface samples the library stores for a scene: those of the bodies
    that emit them (a mesh body's: its vertices)."""
    p = params if params is not None else scene.params
    return int(sum(len(scene.meshes[int(scene.dims[k][0])][0]) if t == SHAPE_MESH
                   else sample_count(int(t), scene.dims[k], p.sample_spacing)
                   for k, t in enumerate(scene.shape_type) if scene.emit[k]))


def scene_footprint(scene, params=None):
    """Estimated HBM bytes a scene adds to a shard: per body its FP32 SDF
    grid, state, flags and solver rows (below a PCG tolerance of 1e-9 also the
    288 entries of its group's inverse blocks and the grouping's integers),
    16 bytes per surface sample (its FP32 position and its share of the
    sample cells), and its expected contacts
    (CONTACTS_PER_BODY per body, BYTES_PER_CONTACT each: contact record and
    solver state)."""
    p = params if params is not None else scene.params
    n = scene.n + 1
    grid = n * (p.grid_n ** 3) * 4
    per_body = 13 * 8 + 64 + 6 * 8 * 12 + 36 * 8 * 18
    if p.pcg_tol < 1e-9:
        per_body += 288 * (8 if p.pcg_precision == PCG_FP64 else 4) + 16 * 4
    return int(grid + 16 * scene_samples(scene, p) + n * per_body + scene.n * CONTACTS_PER_BODY * BYTES_PER_CONTACT)


def _padding(count, params):
    """`count` inert bodies: static tiny boxes far away, emitting no samples."""
    shape = np.zeros(count, np.int32)
    dims = np.full((count, 3), PAD_HALF)
    mass = np.ones(count)
    inertia = np.ones((count, 3))
    static = np.ones(count, np.int32)
    emit = np.zeros(count, np.int32)
    pos = np.zeros((count, 3))
    pos[:, 0] = PAD_FAR + np.arange(count)
    pos[:, 1:] = PAD_FAR
    quat = np.zeros((count, 4))
    quat[:, 3] = 1.0
    z = np.zeros((count, 3))
    return shape, dims, mass, inertia, static, emit, pos, quat, z, z.copy()


class Request:
    __slots__ = ("scene", "arrival", "start", "done")

    def __init__(self, scene, arrival):
        self.scene = scene      # a Resident handle, or an index into SceneServer.resident
        self.arrival = arrival
        self.start = None
        self.done = None


def poisson_requests(rate, duration, n_scenes, seed=0):
    """Arrival times (s) of a Poisson stream, each for a uniformly drawn scene index."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        out.append(Request(int(rng.integers(n_scenes)), t))
    return out


class Resident:
    """Handle of a scene given to the server."""

    def __init__(self, scene, uid):
        self.scene = scene
        self.uid = uid
        self.shard = None
        self.segment = None      # index of its segment in the shard (= scene index in the context)
        self.state = "waiting"   # waiting -> resident -> leaving -> left
        self.steps = 0
        self.footprint = 0

    @property
    def name(self):
        return self.scene.name


class _Segment:
    __slots__ = ("first", "cap", "samples", "owner")

    def __init__(self, first, cap, samples, owner):
        self.first, self.cap, self.samples, self.owner = first, cap, samples, owner


class _Shard:
    """One context, its segments and its worker thread.  Every call on the
    context happens on the worker thread."""

    def __init__(self, server, index, params, budget, device):
        self.server = server
        self.index = index
        self.params = params
        self.budget = budget
        self.ctx = pdc.Context(params, device)
        self.segments = []
        self.pending = collections.deque()
        self.cmds = collections.deque()
        self.cv = threading.Condition()
        self.stop = False
        self.error = None
        self.bytes = 0            # device bytes of the context after the last placement
        self.used = 0             # footprints of the scenes resident or being placed here
        self.rounds = 0
        self.thread = threading.Thread(target=self._loop, name="pdc-shard-%d" % index, daemon=True)
        self.thread.start()

    # ---- called with the server lock held ----
    def room(self):
        return self.budget - self.used

    def post(self, cmd):
        with self.cv:
            self.cmds.append(cmd)
            self.cv.notify()

    def submit(self, req):
        with self.cv:
            self.pending.append(req)
            self.cv.notify()

    # ---- worker ----
    def _loop(self):
        try:
            while True:
                with self.cv:
                    while not (self.stop or self.cmds or self.pending):
                        self.cv.wait()
                    if self.stop and not self.cmds and not self.pending:
                        return
                    cmds = list(self.cmds)
                    self.cmds.clear()
                for cmd in cmds:
                    cmd()
                batch = []
                with self.cv:
                    if self.pending:
                        rest, seen = collections.deque(), set()
                        for r in self.pending:
                            h = r.scene
                            if h.state not in ("resident", "leaving") or h.uid in seen:
                                rest.append(r)
                            else:
                                seen.add(h.uid)
                                batch.append(r)
                        self.pending = rest
                if batch:
                    t0 = time.perf_counter()
                    self.ctx.step_scenes(1, [r.scene.segment for r in batch])
                    t1 = time.perf_counter()
                    self.rounds += 1
                    for r in batch:
                        r.scene.steps += 1
                        r.start = t0 - self.server.t0
                        r.done = t1 - self.server.t0
                    self.server._served(batch)
                self._finish_departures()
        except Exception as exc:  # surfaced by the server's calls
            self.error = exc
            self.server._wake()

    def _finish_departures(self):
        """Segments of leaving scenes with no pending request become free."""
        freed = False
        with self.cv:
            busy = {r.scene.uid for r in self.pending}
        with self.server.lock:
            for seg in self.segments:
                h = seg.owner
                if h is not None and h.state == "leaving" and h.uid not in busy:
                    h.state = "left"
                    seg.owner = None
                    self.used -= h.footprint
                    freed = True
        if freed:
            self.server._admit_waiting()

    def free_segment(self, scene):
        """Index of the smallest free segment that holds the scene (its bodies,
        padding bodies and their samples), or None."""
        n = scene.n
        m = scene_samples(scene, self.params)
        best = None
        for k, seg in enumerate(self.segments):
            if seg.owner is None and seg.cap >= n + 1 and seg.samples >= m:
                if best is None or (seg.cap, seg.samples) < (self.segments[best].cap, self.segments[best].samples):
                    best = k
        return best

    def place(self, h):
        """Worker: writes the scene into a free segment that fits, or appends it."""
        sc = h.scene
        p = self.params
        n = sc.n
        m = scene_samples(sc, p)
        with self.server.lock:
            best = self.free_segment(sc)
            if best is not None:
                self.segments[best].owner = h  # reserved
        fields = (sc.shape_type, self.ctx.mesh_dims(sc), sc.mass, sc.inertia, sc.is_static, sc.emit, sc.pos,
                  sc.quat, sc.vel, sc.angvel)
        if best is not None:
            seg = self.segments[best]
            arrays = [np.concatenate([a, b]) for a, b in zip(fields, _padding(seg.cap - n, p))]
            self.ctx.replace_bodies(seg.first, *arrays[:6], pos=arrays[6], quat=arrays[7], vel=arrays[8],
                                    angvel=arrays[9])
            h.segment = best
        else:
            arrays = [np.concatenate([a, b]) for a, b in zip(fields, _padding(1, p))]
            first = self.ctx.append_bodies(*arrays[:6], pos=arrays[6], quat=arrays[7], vel=arrays[8],
                                           angvel=arrays[9])
            with self.server.lock:
                self.segments.append(_Segment(first, n + 1, m, h))
                h.segment = len(self.segments) - 1
        self.ctx.set_scene_friction(np.array([
            (seg.owner.scene.params.friction_coeff if seg.owner is not None else p.friction_coeff)
            for seg in self.segments], dtype=np.float64))
        with self.server.lock:
            self.bytes = self.ctx.device_bytes()
            h.state = "resident"
        self.server._wake()

    def read_state(self, h, out):
        seg = self.segments[h.segment]
        out.append(self.ctx.get_state_range(seg.first, h.scene.n))
        self.server._wake()

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify()
        self.thread.join()
        self.ctx.close()


class SceneServer:
    """Continuous-batching server over device-resident scenes (module
    docstring).  `params`: the solver parameters every resident scene runs
    with (friction is per scene); `hbm_budget`: device bytes the server may
    hold, shared evenly by its shards."""

    def __init__(self, params, hbm_budget=200 * 2 ** 30, shards=2, max_resident=4096, device=0,
                 reserve_bodies=0, reserve_samples=0):
        self.params = params
        self.max_resident = max_resident
        self.lock = threading.RLock()
        self.done_cv = threading.Condition(self.lock)
        self.t0 = time.perf_counter()
        self.shards = [_Shard(self, k, params, hbm_budget // shards, device) for k in range(shards)]
        if reserve_bodies or reserve_samples:
            for sh in self.shards:
                sh.post(lambda sh=sh: sh.ctx.reserve_bodies(reserve_bodies // shards, reserve_samples // shards))
        self.handles = []
        self.waiting = collections.deque()
        self.served_count = 0
        self.resident = []         # handles in admission order (request indices of run())

    # ---- worker callbacks ----
    def _wake(self):
        with self.lock:
            self.done_cv.notify_all()

    def _served(self, batch):
        with self.lock:
            self.served_count += len(batch)
            self.done_cv.notify_all()

    def _check(self):
        for sh in self.shards:
            if sh.error is not None:
                raise sh.error

    # ---- admission ----
    def n_resident(self):
        return sum(1 for h in self.handles if h.state in ("resident", "leaving", "placing"))

    def _try_admit(self, h):
        """With the lock held: places h on the shard with the most room if it
        fits the budget and the resident limit."""
        if self.n_resident() >= self.max_resident:
            return False
        est = scene_footprint(h.scene, self.params)
        fits = [sh for sh in self.shards if sh.room() >= est and
                (sh.bytes < sh.budget or sh.free_segment(h.scene) is not None)]
        if not fits:
            return False
        sh = max(fits, key=lambda s: s.room())
        sh.used += est
        h.footprint = est
        h.state = "placing"
        h.shard = sh
        sh.post(lambda sh=sh, h=h: sh.place(h))
        return True

    def _admit_waiting(self):
        with self.lock:
            while self.waiting and self._try_admit(self.waiting[0]):
                self.resident.append(self.waiting.popleft())

    def admit(self, scene):
        """Gives a scene to the server: resident at once if it fits, else
        waiting (admitted in order as room frees).  Returns its handle."""
        with self.lock:
            h = Resident(scene, len(self.handles))
            self.handles.append(h)
            if self.waiting or not self._try_admit(h):
                self.waiting.append(h)
            else:
                self.resident.append(h)
            return h

    def admit_all(self, scenes):
        out = [self.admit(sc) for sc in scenes]
        self.wait_placed()
        return out

    def wait_placed(self, timeout=600.0):
        """Waits until no admitted scene is still being written to the device."""
        end = time.perf_counter() + timeout
        with self.lock:
            while any(h.state == "placing" for h in self.handles):
                self._check()
                left = end - time.perf_counter()
                if left <= 0:
                    raise TimeoutError("scenes not placed in time")
                self.done_cv.wait(min(left, 0.5))
            self._check()

    def release(self, h):
        """The scene leaves once its pending requests are served; its segment
        then takes waiting scenes."""
        with self.lock:
            if h.state == "waiting":
                self.waiting.remove(h)
                h.state = "left"
                return
            if h.state == "placing":
                self.wait_placed()
            h.state = "leaving"
        h.shard.post(lambda: None)  # wakes the worker to free the segment

    # ---- requests ----
    def submit(self, h, arrival=0.0):
        r = Request(h, arrival)
        with self.lock:
            if h.state not in ("resident", "placing", "leaving"):
                raise RuntimeError("scene %s is not resident (%s)" % (h.name, h.state))
            sh = h.shard
        sh.submit(r)
        return r

    def drain(self, requests, timeout=600.0):
        end = time.perf_counter() + timeout
        with self.lock:
            while any(r.done is None for r in requests):
                self._check()
                left = end - time.perf_counter()
                if left <= 0:
                    raise TimeoutError("requests not served in time")
                self.done_cv.wait(min(left, 0.5))

    def state(self, h):
        """(pos, quat, vel, angvel) of a resident scene's bodies."""
        self.wait_placed()
        out = []
        h.shard.post(lambda: h.shard.read_state(h, out))
        with self.lock:
            while not out:
                self._check()
                self.done_cv.wait(0.5)
        return out[0]

    def device_bytes(self):
        """Device bytes of the shards' contexts after their last placement."""
        return sum(sh.bytes for sh in self.shards)

    def footprint(self):
        """Footprints held by the resident scenes."""
        return sum(sh.used for sh in self.shards)

    def run(self, requests, time_limit=None):
        """Replays requests (Request with `scene` a handle or an index into
        `resident`, arrival times in seconds from the call): each is handed to
        its scene's shard at its arrival time, while the shards' rounds run.
        Returns the requests served (those not served by the time limit are
        dropped)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if not isinstance(r.scene, Resident):
                r.scene = self.resident[r.scene]
        start = time.perf_counter()
        shift = start - self.t0
        for r in reqs:
            r.start = r.done = None
        i = 0
        while i < len(reqs):
            now = time.perf_counter() - start
            if time_limit is not None and now > time_limit:
                break
            while i < len(reqs) and reqs[i].arrival <= now:
                r = reqs[i]
                r.arrival += shift
                r.scene.shard.submit(r)
                i += 1
            if i < len(reqs):
                time.sleep(max(0.0, min(reqs[i].arrival - (time.perf_counter() - start), 5e-4)))
        sent = reqs[:i]
        end = None if time_limit is None else start + time_limit * 2
        with self.lock:
            while any(r.done is None for r in sent):
                self._check()
                if end is not None and time.perf_counter() > end:
                    break
                self.done_cv.wait(0.05)
        return [r for r in sent if r.done is not None]

    def run_plain(self, requests, time_limit=None):
        """The plain per-scene path: one request per step call, first come
        first served, each waited for before the next (no batching)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if not isinstance(r.scene, Resident):
                r.scene = self.resident[r.scene]
        start = time.perf_counter()
        shift = start - self.t0
        served = []
        for r in reqs:
            now = time.perf_counter() - start
            if time_limit is not None and now > time_limit:
                break
            if r.arrival > now:
                time.sleep(r.arrival - now)
            r.arrival += shift
            r.scene.shard.submit(r)
            self.drain([r])
            served.append(r)
        return served

    def close(self):
        for sh in self.shards:
            sh.close()


def latency_summary(served):
    lat = np.array([r.done - r.arrival for r in served])
    if lat.size == 0:
        return {"requests": 0}
    span = max(r.done for r in served) - min(r.arrival for r in served)
    return {
        "requests": int(lat.size),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "scene_steps_per_s": float(lat.size / span) if span > 0 else None,
    }
