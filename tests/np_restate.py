"""A second, dense numpy statement of the first Newton iteration, written from
the algebra of the primal-dual method rather than from either implementation:

    u  = J v,   J = [-T0^T (at body a), -T1^T (at body b)]           per contact
    ru = M (v - v_old) - J^T f                        (static rows 0)
    rf = -u + (lam0, -lam1 g) - K e0 [- kappa mbar f_n e0]
    rl = lam o c + mu_c,  c = (-f_n, sqrt(|f_t|^2+P) - mu f_n - sqrt(P) - SH)
    W  = diag(1/(eps + l_n), Kt^-1),  l_n = -lam0/(c0 - eps lam0) [+ kappa mbar],
    Kt = eps I - lam1/(c1 - eps lam1) g g^T + lam1 H
    sf = rf + (-rl0/(c0 - eps lam0), g rl1/(c1 - eps lam1))
    A  = M + eps I + P (J^T W J) P,   rhs = P (J^T W sf - ru)        (P: active rows)

The bracketed terms are the spring law (spring_k k, spring_damp d as a
fraction of k): beta = dt k / (dt k + d k), kappa = 1 / (dt (dt k + d k)),
K = -beta depth / (dt s); without it K = e u_n(v0) (restitution e).

Used by the tests to pin the oracle's setup, residual and assembly.
"""
import numpy as np

PERT = 1e-16
SHIFT = 1e-8


def quat_rot(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def constraint(mu, f):
    return np.sqrt(f[1] ** 2 + f[2] ** 2 + PERT) - mu * f[0] - np.sqrt(PERT) - SHIFT


def gradient(f):
    r = np.sqrt(f[1] ** 2 + f[2] ** 2 + PERT)
    return np.array([f[1], f[2]]) / r


def hessian(f):
    p = 1e-6
    z = np.array([f[2], -f[1]])
    return (np.outer(z, z) + p * np.eye(2)) * (f[1] ** 2 + f[2] ** 2 + p) ** -1.5


def radius(mu, fn):
    k = mu * fn + np.sqrt(PERT) + SHIFT
    return np.sqrt(k * k - PERT)


def first_system(ids, T, mass, I0, quat, vel, ang, is_static, dt, g, mu, restitution=0.0, Hu=1e-4,
                 spring=None, depth=None):
    """ids (C,2) global; T (C,2,3,6) [contact][side][col][row].  Returns A, rhs in
    the local numbering of first appearance, and that numbering.  spring:
    (spring_k, spring_damp) for the spring law (then depth (C,) is needed)."""
    C = ids.shape[0]
    order = []
    for i in range(C):
        for j in range(2):
            if ids[i, j] not in order:
                order.append(int(ids[i, j]))
    g2l = {b: k for k, b in enumerate(order)}
    N = len(order)
    act = np.array([not is_static[b] for b in order])
    g = np.asarray(g, dtype=float)
    # mass matrices
    M = np.zeros((6 * N, 6 * N))
    Minv = []
    trace_sum, nact = 0.0, 0
    for k, b in enumerate(order):
        R = quat_rot(quat[b])
        Iw = R @ np.diag(I0[b]) @ R.T
        Mi = np.zeros((6, 6))
        Mi[:3, :3] = np.eye(3) / mass[b]
        Mi[3:, 3:] = R @ np.diag(1.0 / I0[b]) @ R.T
        Minv.append(Mi)
        if act[k]:
            trace_sum += 3 * mass[b] + np.trace(Iw)
            nact += 1
    mbar = trace_sum / nact / 6 if nact else 1.0
    for k, b in enumerate(order):
        if act[k]:
            R = quat_rot(quat[b])
            M[6 * k:6 * k + 3, 6 * k:6 * k + 3] = np.eye(3) * mass[b] / mbar
            M[6 * k + 3:6 * k + 6, 6 * k + 3:6 * k + 6] = R @ np.diag(I0[b]) @ R.T / mbar
        else:
            M[6 * k:6 * k + 6, 6 * k:6 * k + 6] = np.eye(6)
    J = np.zeros((3 * C, 6 * N))
    for i in range(C):
        for j in range(2):
            k = g2l[int(ids[i, j])]
            J[3 * i:3 * i + 3, 6 * k:6 * k + 6] = -T[i, j]      # T[i,j] is (3 cols, 6 rows) -> T^T
    vphys = np.zeros(6 * N)
    for k, b in enumerate(order):
        vphys[6 * k:6 * k + 3] = vel[b]
        vphys[6 * k + 3:6 * k + 6] = ang[b]
    # characteristic speed
    speed = np.zeros(N)
    for i in range(C):
        rel = np.zeros(3)
        has_static = False
        for j in range(2):
            k = g2l[int(ids[i, j])]
            ve = vphys[6 * k:6 * k + 6].copy()
            if act[k]:
                ve[:3] += g * dt
            else:
                has_static = True
            half = -T[i, j] @ ve
            rel += half
            speed[k] = max(speed[k], np.linalg.norm(half))
        if has_static:
            for j in range(2):
                k = g2l[int(ids[i, j])]
                speed[k] = max(speed[k], np.linalg.norm(rel))
    m_act = np.array([mass[b] for b in order])[act]
    s = float(np.sum(speed[act] * m_act) / np.sum(m_act))
    s = max(s, np.linalg.norm(g) * dt * 0.25, 1e-4)
    v = vphys / s
    vold = vphys.copy()
    for k in range(N):
        if act[k]:
            vold[6 * k:6 * k + 3] += g * dt
    vold = vold / s
    if spring is not None:
        k_spring, damp_frac = spring
        tk, damp = dt * k_spring, damp_frac * k_spring
        beta = tk / (tk + damp)
        kappa = 1.0 / (dt * (tk + damp))
        K = -beta * np.asarray(depth, dtype=float) / dt / s
        kmb = kappa * mbar
    else:
        K = restitution * (J @ v)[0::3]
        kmb = 0.0
    f = np.zeros(3 * C)
    f[0::3] = 1.0
    lam = np.full(2 * C, 1e-2)
    c = np.zeros(2 * C)
    gr = np.zeros((C, 2))
    H = np.zeros((C, 2, 2))
    for i in range(C):
        fi = f[3 * i:3 * i + 3]
        c[2 * i] = -fi[0]
        c[2 * i + 1] = constraint(mu, fi)
        gr[i] = gradient(fi)
        H[i] = hessian(fi)
    gap = -np.mean(lam * c)
    muc = 0.1 * gap
    Pact = np.zeros(6 * N)
    for k in range(N):
        if act[k]:
            Pact[6 * k:6 * k + 6] = 1.0
    ru = (M @ (v - vold) - J.T @ f) * Pact
    u = J @ v
    rf = -u.copy()
    for i in range(C):
        rf[3 * i] += lam[2 * i] - K[i] - kmb * f[3 * i]
        rf[3 * i + 1:3 * i + 3] -= lam[2 * i + 1] * gr[i]
    rl = lam * c + muc
    # tangential weights and errors
    wrf = rf.copy()
    for i in range(C):
        tv = u[3 * i + 1:3 * i + 3]
        tn = tv / np.linalg.norm(tv) if np.linalg.norm(tv) > 0 else tv
        fw = 0.0
        for j in range(2):
            k = g2l[int(ids[i, j])]
            if not act[k]:
                continue
            hh = T[i, j][1:3].T @ tn
            fw += hh @ Minv[k] @ hh * mbar
        w = min(radius(mu, f[3 * i]) / (np.linalg.norm(tv) + 1e-8) * fw, 1.0)
        wrf[3 * i + 1:3 * i + 3] *= w
    err = ru @ ru + wrf @ wrf + rl @ rl
    eps = Hu * err ** (0.66667 * 0.5)
    W = np.zeros((3 * C, 3 * C))
    sf = rf.copy()
    for i in range(C):
        ic0 = 1.0 / (c[2 * i] - eps * lam[2 * i])
        ic1 = 1.0 / (c[2 * i + 1] - eps * lam[2 * i + 1])
        ln = -lam[2 * i] * ic0 + kmb
        Kt = eps * np.eye(2) - lam[2 * i + 1] * ic1 * np.outer(gr[i], gr[i]) + lam[2 * i + 1] * H[i]
        W[3 * i, 3 * i] = 1.0 / (eps + ln)
        W[3 * i + 1:3 * i + 3, 3 * i + 1:3 * i + 3] = np.linalg.inv(Kt)
        sf[3 * i] += -rl[2 * i] * ic0
        sf[3 * i + 1:3 * i + 3] += gr[i] * rl[2 * i + 1] * ic1
    JWJ = J.T @ W @ J
    A = M + eps * np.eye(6 * N) + (Pact[:, None] * JWJ * Pact[None, :])
    rhs = (J.T @ W @ sf - ru) * Pact
    return A, rhs, order, dict(mbar=mbar, s=s, eps=eps, err=err)


def group_bodies(key, stat, w, pairs):
    """The body groups of the standard PCG form's preconditioner, stated over
    sets: groups are frozensets of body indices, coupling strengths exact
    fractions n_GH^2 / (w_G w_H) (the square of n_GH / sqrt(w_G w_H)).
    Levels of sizes 1, 2, 4 each run up to three rounds in which every
    unmerged group of the level's size picks its strongest unmerged neighbour
    group of that size (ties: smaller key) and mutual picks merge.  Returns
    per body the index of its group's member with the smallest key."""
    from fractions import Fraction
    n = len(key)
    between = {}
    for a, b, c in pairs:
        if a != b and not stat[a] and not stat[b]:
            e = frozenset((a, b))
            between[e] = between.get(e, 0) + c
    groups = {frozenset([r]) for r in range(n)}
    gkey = lambda G: min(key[r] for r in G)
    wsum = lambda G: sum(w[r] for r in G)

    def contacts(G, H):
        return sum(c for e, c in between.items() if len(e & G) == 1 and len(e & H) == 1 and e <= G | H)

    for size in (1, 2, 4):
        for _ in range(3):
            level = [G for G in groups if len(G) == size and not any(stat[r] for r in G)]
            pick = {}
            for G in level:
                best = None
                for H in level:
                    if H is G:
                        continue
                    c = contacts(G, H)
                    if c == 0:
                        continue
                    s = Fraction(c * c, wsum(G) * wsum(H))
                    if best is None or s > best[0] or (s == best[0] and gkey(H) < gkey(best[1])):
                        best = (s, H)
                if best:
                    pick[G] = best[1]
            merges = [(G, H) for G, H in pick.items() if pick.get(H) is G and gkey(G) < gkey(H)]
            for G, H in merges:
                groups.discard(G)
                groups.discard(H)
                groups.add(G | H)
            if not merges:
                break
    out = np.zeros(n, dtype=np.int64)
    for G in groups:
        holder = min(G, key=lambda r: key[r])
        for r in G:
            out[r] = holder
    return out
