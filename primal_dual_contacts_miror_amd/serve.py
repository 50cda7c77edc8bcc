"""Served scene stepping: a continuous-batching request loop over
device-resident scenes.

No counterpart in the reference, which runs one scene per process.  Scenes
are made resident (uploaded once, state kept in HBM) in request order of
admission while their HBM footprint fits the budget.  A request asks for one
timestep of one resident scene.  The loop runs rounds back to back: a round
takes every pending request (at most one per scene and at most `max_slots`
scenes; a second request for the same scene waits for the next round) and
steps those scenes together with pdc_step_scenes.  Requests that arrive while
a round is on the device join the next round.  A scene's result does not
depend on which other scenes share its round.
"""
import collections
import time

import numpy as np

from . import pdc


def scene_footprint(scene, params=None):
    """HBM bytes a resident scene holds: per body its FP32 SDF grid, surface
    samples (3 doubles each), state and solver rows; per expected contact its
    record and solver state (contacts bounded by the sample count)."""
    p = params if params is not None else scene.params
    n = scene.n
    grid = n * (p.grid_n ** 3) * 4
    area = 0.0
    for k in range(n):
        d = scene.dims[k]
        area += 4 * np.pi * d[0] ** 2 if scene.shape_type[k] == 1 else 8 * (d[0] * d[1] + d[1] * d[2] + d[0] * d[2])
    samples = int(area / (p.sample_spacing ** 2)) + 8 * n
    per_body = 3 * 8 + 4 * 8 + 6 * 8 * 3 + 36 * 8 * 2 + 64
    per_contact = 36 * 8 + 8 * 40
    return grid + samples * 24 + n * per_body + samples * per_contact


class Request:
    __slots__ = ("scene", "arrival", "start", "done")

    def __init__(self, scene, arrival):
        self.scene = scene
        self.arrival = arrival
        self.start = None
        self.done = None


def poisson_requests(rate, duration, n_scenes, seed=0):
    """Arrival times (s) of a Poisson stream and a uniformly drawn resident scene each."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        out.append(Request(int(rng.integers(n_scenes)), t))
    return out


class SceneServer:
    def __init__(self, scenes, hbm_budget=200 * 2 ** 30, max_slots=4096, device=0):
        self.max_slots = max_slots
        self.resident = []      # indices into `scenes` admitted
        self.rejected = []
        used = 0
        for k, sc in enumerate(scenes):
            b = scene_footprint(sc)
            if used + b <= hbm_budget and len(self.resident) < max_slots:
                self.resident.append(k)
                used += b
            else:
                self.rejected.append(k)
        self.footprint = used
        self.scenes = [scenes[k] for k in self.resident]
        self.system = pdc.BatchedSystem(self.scenes, device=device)

    @classmethod
    def over(cls, system, max_slots=4096):
        """A server over an existing BatchedSystem: all its scenes are resident."""
        srv = cls.__new__(cls)
        srv.max_slots = max_slots
        srv.system = system
        srv.scenes = list(system.scenes)
        srv.resident = list(range(len(srv.scenes)))
        srv.rejected = []
        srv.footprint = sum(scene_footprint(sc) for sc in srv.scenes)
        return srv

    def close(self):
        self.system.ctx.close()

    def run(self, requests, batched=True, time_limit=None):
        """Serves the requests (arrival times relative to the start of the call).
        With batched=False every request is stepped alone, first come first
        served (the plain per-scene path).  Returns the served requests."""
        pending = collections.deque()
        reqs = sorted(requests, key=lambda r: r.arrival)
        i = 0
        served = []
        t0 = time.perf_counter()
        while i < len(reqs) or pending:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            while i < len(reqs) and reqs[i].arrival <= now:
                pending.append(reqs[i])
                i += 1
            if not pending:
                time.sleep(max(0.0, min(reqs[i].arrival - now, 1e-3)))
                continue
            if batched:
                batch, rest, seen = [], collections.deque(), set()
                for r in pending:
                    if r.scene in seen or len(batch) >= self.max_slots:
                        rest.append(r)
                    else:
                        seen.add(r.scene)
                        batch.append(r)
                pending = rest
            else:
                batch = [pending.popleft()]
            start = time.perf_counter() - t0
            self.system.step(1, [r.scene for r in batch])
            done = time.perf_counter() - t0
            for r in batch:
                r.start = start
                r.done = done
            served.extend(batch)
        return served


def latency_summary(served, duration=None):
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
