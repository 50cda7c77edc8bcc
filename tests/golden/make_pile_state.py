"""Fixture generator: the settled state of a seeded pile, i.e. the state after
the workload's own step count (pile_2000: 1000 steps of dt 0.005) from the
seeded drop, computed by the oracle (CPU FP64 restatement).  Stored as .npz
with pos (N,3), quat (N,4), vel (N,3), angvel (N,3).  Used as an identical
input state for per-step parity at ~30k contacts.

    python tests/golden/make_pile_state.py pile_2000 tests/golden/pile_2000_settled.npz [steps]
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

from primal_dual_contacts_miror_amd import scenes  # noqa: E402


def main():
    name = sys.argv[1]
    out = sys.argv[2]
    sc = scenes.WORKLOADS[name]()
    steps = int(sys.argv[3]) if len(sys.argv) > 3 else sc.steps
    from oracle import orc
    p = sc.params
    p.pcg_tol = 1e-8
    o = orc.OracleScene(sc, p)
    t0 = time.time()
    for k in range(steps):
        st = o.step(1)
        if k % 50 == 0:
            print(k, st.n_contacts, st.newton_iterations, "%.1fs" % (time.time() - t0), flush=True)
    pos, quat, vel, ang = o.get_state()
    np.savez(out, pos=pos, quat=quat, vel=vel, angvel=ang, steps=steps, name=name)


if __name__ == "__main__":
    main()
