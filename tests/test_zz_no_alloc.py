"""GPU: pdc_fill_forces_noalloc, the entry point that allocates no device
memory (its buffers come from the library's fixed scratch array), against
pdc_fill_forces on the same inputs: bit-identical increments, local forces
and statistics, on two calls in a row, the second with fewer contacts; the
context stays usable by the allocating entry point afterwards; requests
beyond the scratch limits are refused with PDC_ERR_CAPACITY."""
import os

import numpy as np
import pytest

from primal_dual_contacts_miror_amd import pdc, scenes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _same(a, b):
    acc_a, lf_a, st_a = a
    acc_b, lf_b, st_b = b
    np.testing.assert_array_equal(acc_a, acc_b)
    np.testing.assert_array_equal(lf_a, lf_b)
    assert st_a.as_dict() == st_b.as_dict()


def test_fill_forces_noalloc_equals_fill_forces(require_gpu):
    z = np.load(os.path.join(HERE, "golden", "pile_2000_settled.npz"))
    sc = scenes.pile_2000()
    state = (z["pos"], z["quat"], z["vel"], z["angvel"])
    ds = pdc.DynamicSystem(sc)
    ds.set_state(*state)
    ids, pos, depth, T, lf = ds.contacts()
    ds.ctx.close()
    quat, vel, ang = state[1], state[2], state[3]
    print("settled pile: %d contacts" % len(ids))
    assert 20000 < len(ids) <= 32768
    noalloc = pdc.PrimalDualForceSolver(sc)   # fresh context: its first call is the no-allocation one
    plain = pdc.PrimalDualForceSolver(sc)
    for n in (len(ids), 4000):                # two calls in a row, the second smaller
        got = noalloc.ctx.fill_forces(ids[:n], T[:n], depth[:n], quat, vel, ang, noalloc=True)
        want = plain.ctx.fill_forces(ids[:n], T[:n], depth[:n], quat, vel, ang)
        print("%d contacts: Newton %d, PCG %d" % (n, got[2].newton_iterations, got[2].pcg_iterations))
        assert got[2].n_contacts == n and got[2].newton_iterations > 0
        _same(got, want)
    # the allocating entry point on the same context afterwards
    _same(noalloc.ctx.fill_forces(ids[:2000], T[:2000], depth[:2000], quat, vel, ang),
          plain.ctx.fill_forces(ids[:2000], T[:2000], depth[:2000], quat, vel, ang))
    # beyond the contact limit: refused before any device work
    big = np.tile(ids, (2, 1))[:40000]
    with pytest.raises(pdc.PdcError) as e:
        noalloc.ctx.fill_forces(big, np.tile(T, (2, 1))[:40000], np.tile(depth, 2)[:40000], quat, vel, ang,
                                noalloc=True)
    assert e.value.code == pdc.PDC_ERR_CAPACITY
    noalloc.ctx.close()
    plain.ctx.close()
