"""GPU: the body groups of the standard PCG form's preconditioner, as the
device builds them (pdc_get_groups), equal the oracle's group_bodies on the
same contact list: keys the global ids, weights the bodies' contact counts,
one pair per two bodies in contact.  On a single scene through fillForces
(the 210-brick pyramid and the settled 2000-body pile) and on a batch of 64
random piles stepped into contact (groups never cross scenes)."""
import os

import numpy as np
import pytest

from oracle import orc
from primal_dual_contacts_miror_amd import pdc, scenes
from primal_dual_contacts_miror_amd.params import PCG_FP64

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TIGHT = dict(pcg_precision=PCG_FP64, pcg_tol=1e-12)


def oracle_groups(n_bodies, is_static, ids):
    """group_bodies over the bodies in the contact list `ids` (C, 2), as
    global keys per body (-1 outside the list)."""
    bodies = np.unique(ids)
    loc = {int(b): k for k, b in enumerate(bodies)}
    w = np.zeros(len(bodies), np.int64)
    pairs = {}
    for a, b in ids:
        w[loc[int(a)]] += 1
        w[loc[int(b)]] += 1
        key = (min(loc[int(a)], loc[int(b)]), max(loc[int(a)], loc[int(b)]))
        pairs[key] = pairs.get(key, 0) + 1
    rep = orc.group_bodies(bodies.astype(np.int64), np.asarray(is_static)[bodies], w,
                           [(a, b, c) for (a, b), c in pairs.items()])
    out = np.full(n_bodies, -1, np.int32)
    out[bodies] = bodies[rep]
    return out


def _summary(g):
    keys, sizes = np.unique(g[g >= 0], return_counts=True)
    return "%d bodies in %d groups, sizes %s" % ((g >= 0).sum(), len(keys), np.bincount(sizes)[1:].tolist())


@pytest.mark.parametrize("which", ["pyramid_210", "pile_2000"])
def test_groups_equal_oracle_fill_forces(require_gpu, which):
    if which == "pyramid_210":
        sc = scenes.pyramid_210(**TIGHT)
        o = orc.OracleScene(sc)
        o.set_threads(16)
        o.step(25)
        state = o.get_state()
    else:
        sc = scenes.pile_2000(**TIGHT)
        z = np.load(os.path.join(HERE, "golden", "pile_2000_settled.npz"))
        state = (z["pos"], z["quat"], z["vel"], z["angvel"])
    ds = pdc.DynamicSystem(sc, sc.params)
    ds.set_state(*state)
    ids, pos, depth, T, lf = ds.contacts()
    ds.ctx.close()
    solver = pdc.PrimalDualForceSolver(sc, sc.params)
    co = orc.Contacts(ids, pos, depth, T)
    solver.fill_forces(co, state[1], state[2], state[3])
    got = solver.ctx.groups()
    solver.ctx.close()
    want = oracle_groups(sc.n, sc.is_static, ids)
    print(which, _summary(got))
    assert np.array_equal(got, want)
    assert (got >= 0).sum() > 0 and len(np.unique(got[got >= 0])) < (got >= 0).sum()


def test_groups_equal_oracle_batch(require_gpu):
    scs = [scenes.random_pile(1000 + k, **TIGHT) for k in range(64)]
    bs = pdc.BatchedSystem(scs)
    for _ in range(120):
        st = bs.step(1)
    ids = bs.ctx.get_contacts(sum(x.n_contacts for x in st))[0]  # the last step's contacts
    got = bs.ctx.groups()
    n = sum(s.n for s in scs)
    stat = np.concatenate([s.is_static for s in scs])
    want = oracle_groups(n, stat, ids)
    print("batch of 64:", _summary(got))
    assert np.array_equal(got, want)
    bs.ctx.close()


def test_groups_need_a_grouped_solve(require_gpu):
    sc = scenes.stack_10()  # pcg_tol 1e-8: the single-reduction form, no groups
    ds = pdc.DynamicSystem(sc, sc.params)
    ds.step(1)
    with pytest.raises(pdc.PdcError) as e:
        ds.ctx.groups()
    assert e.value.code == pdc.PDC_ERR_STATE
    ds.ctx.close()
