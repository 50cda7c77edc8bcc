"""Independent scenes stepped together in one context (batched / served mode).

Every scene runs the same pipeline as when it is alone; tiles, reductions and
the Newton decisions are per scene, so a scene's result must be bit-identical
alone and in any mix.  The single-scene path itself is checked against the
oracle in test_gpu_parity.py; one batch here is also checked against the
oracle directly.
"""
import numpy as np
import pytest

from oracle import orc
from primal_dual_contacts_miror_amd import pdc, scenes
from primal_dual_contacts_miror_amd.params import PCG_FP64

pytestmark = pytest.mark.gpu


def _piles(seeds, n_max=30, **kw):
    """Small seeded piles dropped from a low slab under strong gravity, so the
    bodies are in contact after a few dozen steps."""
    out = []
    for s in seeds:
        rng = np.random.default_rng(s)
        n = int(rng.integers(8, n_max + 1))
        sc = scenes.pile(n, seed=s, half=3.0, height=3.6, name="low_pile_%d" % n, grid_n=32, **kw)
        sc.params.gravity[1] = -40.0
        out.append(sc)
    return out


STEPS = 100


def _alone(sc, steps, params=None):
    ds = pdc.DynamicSystem(sc, params if params is not None else sc.params)
    stats = None
    contacts = 0
    for _ in range(steps):
        stats = ds.step(1)
        contacts += stats.n_contacts
    st = ds.state()
    ds.ctx.close()
    assert contacts > 0, "the scene never came into contact"
    return st, stats


def _same_state(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_batch_equals_each_scene_alone(require_gpu):
    scs = _piles([11, 12, 13, 14, 15])
    bs = pdc.BatchedSystem(scs)
    last = None
    newton = 0
    for _ in range(STEPS):
        last = bs.step(1)
        newton += sum(s.newton_iterations for s in last)
    assert newton > 0 and sum(s.n_contacts for s in last) > 0
    for k, sc in enumerate(scs):
        alone, st = _alone(sc, STEPS)
        assert _same_state(bs.state(k), alone), "scene %d differs in the batch" % k
        assert last[k].n_contacts == st.n_contacts
        assert last[k].newton_iterations == st.newton_iterations
        assert last[k].pcg_iterations == st.pcg_iterations
        assert last[k].line_search_trials == st.line_search_trials
    bs.ctx.close()


def test_subset_steps_only_selected_scenes(require_gpu):
    scs = _piles([21, 22, 23, 24])
    bs = pdc.BatchedSystem(scs)
    init = [bs.state(k) for k in range(len(scs))]
    sel = [0, 2]
    for _ in range(STEPS):
        st = bs.step(1, sel)
        assert len(st) == len(sel)
    for k, sc in enumerate(scs):
        if k in sel:
            alone, _ = _alone(sc, STEPS)
            assert _same_state(bs.state(k), alone)
        else:
            assert _same_state(bs.state(k), init[k])
    # the other scenes join later and follow their own trajectory, while the
    # first two keep stepping in the mix
    for _ in range(STEPS):
        bs.step(1, [1, 3, 0])
    for k in (1, 3):
        alone, _ = _alone(scs[k], STEPS)
        assert _same_state(bs.state(k), alone)
    alone0, _ = _alone(scs[0], 2 * STEPS)
    assert _same_state(bs.state(0), alone0)
    bs.ctx.close()


def test_per_scene_friction(require_gpu):
    lo, base, hi = (_piles([31], friction_coeff=mu)[0] for mu in (0.2, 0.6, 0.9))
    bs = pdc.BatchedSystem([lo, base, hi])
    for _ in range(STEPS):
        bs.step(1)
    states = []
    for k, sc in enumerate([lo, base, hi]):
        alone, _ = _alone(sc, STEPS)
        assert _same_state(bs.state(k), alone)
        states.append(alone[0])
    assert not np.array_equal(states[0], states[2])  # the friction coefficient matters
    bs.ctx.close()


def test_batch_against_oracle(require_gpu):
    # resting stacks of different heights and friction coefficients: the
    # Newton solve converges to 1e-10, so each scene of the batch must follow
    # the oracle's FP64 trajectory (direct solve) to 1e-6 of the body size
    scs = [scenes.stack_10(n_cubes=n, tolerance=1e-10, pcg_tol=1e-13, friction_coeff=mu)
           for n, mu in ((4, 0.3), (7, 0.5), (10, 0.8))]
    bs = pdc.BatchedSystem(scs)
    steps = 40
    for _ in range(steps):
        st = bs.step(1)
    for k, sc in enumerate(scs):
        o = orc.OracleScene(sc)
        o.set_linear_solver(orc.SOLVER_DIRECT)
        for _ in range(steps):
            so = o.step(1)
        assert so.n_contacts == st[k].n_contacts
        po, qo = o.get_state()[:2]
        pg, qg = bs.state(k)[:2]
        dyn = sc.is_static == 0
        ep = np.max(np.linalg.norm(pg[dyn] - po[dyn], axis=1))  # unit cubes: body size 1
        ea = np.max(2 * np.arccos(np.clip(np.abs(np.sum(qg[dyn] * qo[dyn], axis=1)), 0, 1)))
        print("scene %d: %d cubes, position error %.3e, angle error %.3e rad" % (k, sc.n - 1, ep, ea))
        assert ep < 1e-6 and ea < 1e-6
    bs.ctx.close()


def test_shapes_per_size_class_and_autotune(require_gpu):
    """Scenes of different size classes with non-default work shapes: the
    batch still equals each scene alone under the same shape table; the
    autotuner returns a valid table and leaves the state untouched."""
    from primal_dual_contacts_miror_amd.params import PdcShapes
    big = scenes.pile(90, seed=43, half=4.0, height=6.0, name="pile_90", grid_n=32)
    big.params.gravity[1] = -40.0
    scs = _piles([41, 42]) + [big]   # size classes 0 and 1
    table = PdcShapes()
    for c, (ct, bw) in enumerate(((64, 1), (128, 2), (256, 4), (256, 4))):
        table.contact_tile[c] = ct
        table.body_waves[c] = bw
    table.pcg_chunk = 8
    bs = pdc.BatchedSystem(scs)
    bs.ctx.set_shapes(table)
    for _ in range(40):
        bs.step(1)
    for k, sc in enumerate(scs):
        ds = pdc.DynamicSystem(sc, sc.params)
        ds.ctx.set_shapes(table)
        for _ in range(40):
            ds.step(1)
        assert _same_state(bs.state(k), ds.state()), "scene %d differs under the shape table" % k
        ds.ctx.close()
    before = [bs.state(k) for k in range(len(scs))]
    chosen = bs.ctx.autotune(1)
    d = chosen.as_dict()
    assert all(t in (64, 128, 256) for t in d["contact_tile"])
    assert all(w in (1, 2, 4) for w in d["body_waves"])
    assert d["pcg_chunk"] in (0, 8, 16, 32)
    assert bs.ctx.get_shapes().as_dict() == d
    for k in range(len(scs)):
        assert _same_state(bs.state(k), before[k])
    bs.step(1)  # stepping continues with the chosen table
    bs.ctx.close()
