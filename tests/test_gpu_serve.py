"""Served stepping: continuous batching over device-resident scenes.

Requests that overlap share rounds, rounds of the two shards run on the
device together, scenes leave and others take their segments; every scene's
state after its requests must be bit-identical to the scene stepped alone the
same number of times.  Scenes that do not fit the budget or the resident
limit wait and are admitted as others leave.
"""
import numpy as np
import pytest

from primal_dual_contacts_miror_amd import pdc, serve
from test_gpu_batch import _piles

pytestmark = pytest.mark.gpu


def _alone_state(sc, steps):
    ds = pdc.DynamicSystem(sc, sc.params)
    for _ in range(steps):
        ds.step(1)
    st = ds.state()
    ds.ctx.close()
    return st


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_served_results_equal_alone(require_gpu):
    scs = _piles([51, 52, 53, 54, 55])
    srv = serve.SceneServer(scs[0].params, shards=2)
    hs = srv.admit_all(scs)
    assert all(h.state == "resident" for h in hs)
    assert {h.shard.index for h in hs} == {0, 1}
    rng = np.random.default_rng(3)
    # bursts of overlapping requests, repeats of one scene included; arrivals
    # spread so later ones land while earlier rounds are on the device
    reqs = [serve.Request(int(rng.integers(5)), 0.002 * k) for k in range(300)]
    served = srv.run(reqs)
    assert len(served) == len(reqs)
    rounds = {}
    for r in served:
        rounds.setdefault((r.scene.shard.index, r.done), []).append(r)
    assert max(len(v) for v in rounds.values()) > 1, "no round held more than one scene"
    for h, sc in zip(hs, scs):
        assert h.steps == sum(1 for r in served if r.scene is h)
        assert _same(srv.state(h), _alone_state(sc, h.steps)), "scene %s" % sc.name
    summary = serve.latency_summary(served)
    assert summary["requests"] == len(reqs) and summary["p99_ms"] >= summary["p50_ms"] > 0
    srv.close()


def test_churn_arrive_step_leave_readmit(require_gpu):
    """Scenes arrive, step, leave and come back while others keep stepping:
    a waiting scene is admitted when a resident one leaves, a departed
    scene's segment is reused in place (the scene that comes back, or a
    smaller one), and every scene ends bit-identical to itself stepped
    alone."""
    scs = _piles([71, 73, 74, 75, 76])
    srv = serve.SceneServer(scs[0].params, shards=1, max_resident=3)
    a, b, c, d = [srv.admit(sc) for sc in scs[:4]]
    srv.wait_placed()
    assert [h.state for h in (a, b, c)] == ["resident"] * 3 and d.state == "waiting"
    srv.drain([srv.submit(h) for h in (a, b, c) for _ in range(20)])
    srv.release(b)
    srv.drain([srv.submit(a)])
    srv.wait_placed()
    assert b.state == "left" and d.state == "resident"
    srv.drain([srv.submit(h) for h in (a, c, d) for _ in range(25)])
    # c leaves and comes back as a fresh copy of its scene: the same bodies fit its old segment
    srv.release(c)
    srv.drain([srv.submit(a)])
    c2 = srv.admit(scs[2])
    srv.wait_placed()
    assert c2.state == "resident"
    sh = a.shard
    assert len(sh.segments) < 5, "no departed scene's segment was reused"
    srv.drain([srv.submit(h) for h in (a, d, c2) for _ in range(30)])
    for h, sc in ((a, scs[0]), (d, scs[3]), (c2, scs[2])):
        assert _same(srv.state(h), _alone_state(sc, h.steps)), "scene %s after %d steps" % (sc.name, h.steps)
    assert a.steps == 20 + 1 + 25 + 1 + 30
    srv.close()


def test_admission_by_footprint(require_gpu):
    scs = _piles([61, 62, 61])  # the third is the first one again: it fits the first's room and segment
    sizes = [serve.scene_footprint(s) for s in scs]
    srv = serve.SceneServer(scs[0].params, hbm_budget=sizes[0] + sizes[1] + sizes[2] // 2, shards=1)
    hs = [srv.admit(sc) for sc in scs]
    srv.wait_placed()
    assert [h.state for h in hs] == ["resident", "resident", "waiting"]
    assert srv.footprint() == sizes[0] + sizes[1]
    assert srv.device_bytes() > 0
    served = srv.run([serve.Request(hs[0], 0.0), serve.Request(hs[1], 0.0)])
    assert len(served) == 2
    # the first scene leaves: its footprint frees and the waiting scene is admitted
    srv.release(hs[0])
    srv.drain([srv.submit(hs[1])])
    srv.wait_placed()
    assert hs[0].state == "left" and hs[2].state == "resident"
    assert srv.footprint() == sizes[1] + sizes[2]
    srv.close()
