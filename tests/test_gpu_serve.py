"""Served stepping: continuous batching over device-resident scenes.

Requests that overlap on the device share rounds; every scene's state after
its requests must be bit-identical to the scene stepped alone the same number
of times, and scenes that do not fit the HBM budget are not admitted.
"""
import collections

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


def test_served_results_equal_alone(require_gpu):
    scs = _piles([51, 52, 53, 54, 55])
    srv = serve.SceneServer(scs)
    assert srv.resident == list(range(5))
    rng = np.random.default_rng(3)
    # bursts of overlapping requests, repeats of one scene included; arrivals
    # spread so later ones land while earlier rounds are on the device
    reqs = [serve.Request(int(rng.integers(5)), 0.002 * k) for k in range(300)]
    served = srv.run(reqs)
    assert len(served) == len(reqs)
    counts = collections.Counter(r.scene for r in served)
    rounds = collections.Counter(r.done for r in served)
    assert max(rounds.values()) > 1, "no round held more than one scene"
    for k, sc in enumerate(scs):
        assert np.array_equal(srv.system.state(k)[0], _alone_state(sc, counts[k])[0]), "scene %d" % k
    summary = serve.latency_summary(served)
    assert summary["requests"] == len(reqs) and summary["p99_ms"] >= summary["p50_ms"] > 0
    srv.close()


def test_admission_by_footprint(require_gpu):
    scs = _piles([61, 62, 63])
    sizes = [serve.scene_footprint(s) for s in scs]
    srv = serve.SceneServer(scs, hbm_budget=sizes[0] + sizes[1] + sizes[2] // 2)
    assert srv.resident == [0, 1] and srv.rejected == [2]
    assert srv.footprint == sizes[0] + sizes[1]
    served = srv.run([serve.Request(0, 0.0), serve.Request(1, 0.0)])
    assert len(served) == 2 and served[0].done == served[1].done  # one shared round
    srv.close()
