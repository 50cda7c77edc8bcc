"""The body groups and the group-block preconditioner of the standard PCG
form (pcg_tol < 1e-9), in the oracle.

The grouping is pinned against np_restate.group_bodies, an independent
statement over sets with exact fractions, on random contact graphs, and by
hand on a chain and a star.  The grouped PCG is pinned by what the
mathematics fixes: one group holding every row makes the preconditioner the
exact inverse (one iteration), singleton groups are block Jacobi (the same
bits), and on a system of strongly coupled pairs grouping the pairs needs
fewer iterations for the same solution.
"""
import numpy as np
import pytest

from oracle import orc
import np_restate


def _random_graph(rng, n, m, static_frac=0.1):
    key = rng.permutation(10 * n)[:n].astype(np.int64)
    stat = (rng.random(n) < static_frac).astype(np.int32)
    pairs = {}
    for _ in range(m):
        a, b = rng.integers(n, size=2)
        if a == b:
            continue
        k = (min(a, b), max(a, b))
        pairs[k] = pairs.get(k, 0) + int(rng.integers(1, 6))
    pl = [(a, b, c) for (a, b), c in pairs.items()]
    w = np.zeros(n, np.int64)
    for a, b, c in pl:
        w[a] += c
        w[b] += c
    w += rng.integers(0, 4, size=n)  # contacts with bodies outside the list (static partners)
    w = np.maximum(w, 1)
    return key, stat, w, pl


@pytest.mark.parametrize("seed", range(6))
def test_grouping_matches_set_restatement(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(10, 60))
    key, stat, w, pairs = _random_graph(rng, n, int(rng.integers(n, 4 * n)))
    got = orc.group_bodies(key, stat, w, pairs)
    want = np_restate.group_bodies(key, stat, w, pairs)
    assert np.array_equal(got, want)
    # invariants: sizes 1, 2, 4, 8; static bodies alone; the holder has the
    # smallest key of its group; every group is connected by contacts
    groups = {}
    for r, h in enumerate(got):
        groups.setdefault(h, []).append(r)
    adj = {r: set() for r in range(n)}
    for a, b, _ in pairs:
        adj[a].add(b)
        adj[b].add(a)
    for h, mem in groups.items():
        assert len(mem) in (1, 2, 4, 8)
        assert key[h] == min(key[r] for r in mem)
        if len(mem) > 1:
            assert not any(stat[r] for r in mem)
            seen, todo = {mem[0]}, [mem[0]]
            while todo:
                x = todo.pop()
                for y in adj[x]:
                    if y in mem and y not in seen:
                        seen.add(y)
                        todo.append(y)
            assert seen == set(mem)
    assert len(groups) < n  # something merged on these graphs


def test_grouping_by_hand():
    # chain 0-1-2-3, strong ends and a weak middle: pairs first, then the two pairs
    key = np.array([5, 6, 7, 8])
    pairs = [(0, 1, 5), (1, 2, 1), (2, 3, 5)]
    w = np.array([5, 6, 6, 5])
    assert orc.group_bodies(key, np.zeros(4), w, pairs).tolist() == [0, 0, 0, 0]
    # the weak link broken: two pairs
    assert orc.group_bodies(key, np.zeros(4), w, [pairs[0], pairs[2]]).tolist() == [0, 0, 2, 2]
    # star: the centre pairs with one leaf (equal strengths: the smallest key);
    # the other leaves have no partner left and stay alone
    key = np.array([10, 3, 2, 4])
    pairs = [(0, 1, 2), (0, 2, 2), (0, 3, 2)]
    w = np.array([6, 2, 2, 2])
    assert orc.group_bodies(key, np.zeros(4), w, pairs).tolist() == [2, 1, 2, 3]
    # a static centre groups with nobody
    assert orc.group_bodies(key, np.array([1, 0, 0, 0]), w, pairs).tolist() == [0, 1, 2, 3]
    # the strength is normalised by the partner's contacts: 0 prefers 2 (3^2/4) over 1 (4^2/20)
    key = np.array([0, 1, 2])
    assert orc.group_bodies(key, np.zeros(3), np.array([7, 20, 4]), [(0, 1, 4), (0, 2, 3)]).tolist() == [0, 1, 0]


def _coupled_pairs_system(rng, nb, k_strong=500.0, k_weak=0.5):
    """SPD block system of nb bodies, built as it is in the solver: a mass
    block per body plus, per contact, a positive semidefinite term
    k (J_a - J_b)^T (J_a - J_b).  Bodies 2k and 2k+1 share stiff contacts,
    neighbouring pairs soft ones."""
    n = 6 * nb
    A = np.zeros((n, n))
    for b in range(nb):
        Q = rng.normal(size=(6, 6))
        A[6 * b:6 * b + 6, 6 * b:6 * b + 6] = Q @ Q.T / 6 + np.eye(6)

    def contact(a, b, k):
        J = np.zeros((3, n))
        T = rng.normal(size=(3, 6))
        J[:, 6 * a:6 * a + 6] = T
        J[:, 6 * b:6 * b + 6] = -T
        A[:] += k * J.T @ J

    for b in range(0, nb - 1, 2):
        for _ in range(3):
            contact(b, b + 1, k_strong)
    for b in range(1, nb - 1, 2):
        contact(b, b + 1, k_weak)
    return A


def test_grouped_pcg():
    rng = np.random.default_rng(7)
    nb = 16
    A = _coupled_pairs_system(rng, nb)
    assert np.min(np.linalg.eigvalsh(A)) > 0
    b = rng.normal(size=6 * nb)
    x = np.linalg.solve(A, b)
    # singleton groups are block Jacobi, bit for bit
    xj, itj = orc.pcg_solve(A, b, tol=1e-12, maxit=2000)
    xs, its = orc.pcg_solve(A, b, tol=1e-12, maxit=2000, rep=np.arange(nb))
    assert its == itj and np.array_equal(xs, xj)
    # the strongly coupled pairs grouped: the same solution in fewer iterations
    pairs = np.repeat(np.arange(0, nb, 2), 2)
    xg, itg = orc.pcg_solve(A, b, tol=1e-12, maxit=2000, rep=pairs)
    np.testing.assert_allclose(xg, x, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(xj, x, rtol=1e-9, atol=1e-11)
    assert itg < itj / 2
    # one group of eight bodies holding every row: the exact inverse, one iteration
    A8 = _coupled_pairs_system(rng, 8)
    b8 = rng.normal(size=48)
    x8, it8 = orc.pcg_solve(A8, b8, tol=1e-12, maxit=100, rep=np.zeros(8))
    assert it8 == 1
    np.testing.assert_allclose(x8, np.linalg.solve(A8, b8), rtol=1e-10, atol=1e-12)
    with pytest.raises(ValueError):
        orc.pcg_solve(A8, b8, rep=np.full(8, 9))
