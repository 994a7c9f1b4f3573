"""NumPy emulation of the multifrontal kernels' data flow (CPU test helper).

Executes, front by front and level by level, the same operations the HIP kernels of
``sawsim_amd/csrc/mf.hip`` perform, on a flat arena with the same offsets, the same
extend-add maps, entry offsets and blocked panel order (diagonal LU, panel solves,
L-shaped update, Schur complement), including the storage convention of the diagonal
blocks: after the factorization each nbl x nbl diagonal block holds L_D^{-1} strictly below
its diagonal (unit diagonal implied) and U_D^{-1} on and above it, nbl being the level's
panel width (32 for levels whose largest pivot block is <= 128 when nb is 64, else nb).
It checks the host plan arrays and the blocked algorithm on the CPU; it is not used by the
product path.
"""
import numpy as np


def panel_width(nb, level_pmax):
    return 32 if (nb == 64 and level_pmax <= 128) else nb


def _level_widths(mf, nb):
    out = []
    for L in range(mf.lvl_ptr.size - 1):
        fronts = mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]
        out.append(panel_width(nb, int(mf.p[fronts].max())))
    return out


def row_scales(mf, Kaa, Maa, Daa):
    """Row maxima of |K|, |M|, |D| over the active entries (the pivot guard's row scale
    s_g(w) = |K|_g + w^2 |M|_g + w |beta| |D|_g)."""
    n = int(mf.n)
    f = mf.ent_front
    g = mf.rows[mf.rows_ptr[f] + mf.ent_loc % mf.m[f]]
    out = []
    for X in (Kaa, Maa, Daa):
        r = np.zeros(n)
        if X is not None:
            np.maximum.at(r, g, np.abs(X[mf.ent_slot]))
        out.append(r)
    return out


def guard_pivot(piv, scale, eps):
    """The kernels' pivot guard: |piv| < eps * scale (or non-finite) -> eps * scale with the
    pivot's phase (eps * scale for a zero pivot), eps = sqrt(unit roundoff) of the front
    precision.  Returns (pivot, replaced)."""
    mag = abs(complex(piv))
    tau = eps * scale
    if mag >= tau and np.isfinite(mag):
        return piv, False
    if 0 < mag < np.inf:
        return piv * (tau / mag), True
    return tau, True


def factor(mf, B, omegas, beta, Kaa, Maa, Daa, nb=32, dtype=np.complex128, stats=None):
    """Factor arena of B frequencies; ``stats`` (a dict, optional) receives the number of pivots
    the guard replaced ('perturbed')."""
    off = mf.off
    arena = np.zeros(mf.arena_elems * B, dtype=dtype)
    widths = _level_widths(mf, nb)
    eps = np.sqrt(2.0 ** -24 if np.dtype(dtype) == np.complex64 else 2.0 ** -53)
    rk, rm, rd = row_scales(mf, Kaa, Maa, Daa)
    nrep = 0

    def front(f, b):
        m = int(mf.m[f])
        base = off[f] * B + b * m * m
        return arena[base:base + m * m].reshape(m, m, order='F')
    for L in range(mf.lvl_ptr.size - 1):
        fronts = mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]
        e0, e1 = mf.ent_lvl_ptr[L], mf.ent_lvl_ptr[L + 1]
        for b in range(B):
            w = omegas[b]
            for f in fronts:
                Fm = front(f, b)
                Fm[:] = 0
                for slot in range(mf.child_ptr[f], mf.child_ptr[f + 1]):
                    ch = mf.child_ids[slot]
                    inv = mf.inv[mf.inv_ptr[slot]:mf.inv_ptr[slot + 1]]
                    Fc = front(ch, b)
                    pc = mf.p[ch]
                    sel = np.flatnonzero(inv >= 0)
                    Fm[np.ix_(sel, sel)] += Fc[pc + inv[sel][:, None], pc + inv[sel][None, :]]
            s = mf.ent_slot[e0:e1]
            a = Kaa[s] - w * w * Maa[s]
            if Daa is not None:
                a = a + 1j * w * beta * Daa[s]
            for k in range(e0, e1):
                f = mf.ent_front[k]
                Fm = front(f, b)
                loc = mf.ent_loc[k]
                m = mf.m[f]
                Fm[loc % m, loc // m] += a[k - e0]
            nbl = widths[L]
            for f in fronts:
                Fm = front(f, b)
                m, p = int(mf.m[f]), int(mf.p[f])
                rows = mf.rows[mf.rows_ptr[f]:mf.rows_ptr[f + 1]]
                for k0 in range(0, p, nbl):
                    k1 = min(p, k0 + nbl)
                    D = Fm[k0:k1, k0:k1]
                    for k in range(k1 - k0):
                        g = rows[k0 + k]
                        D[k, k], rep = guard_pivot(D[k, k], rk[g] + w * w * rm[g] + w * abs(beta) * rd[g], eps)
                        nrep += rep
                        D[k + 1:, k] /= D[k, k]
                        D[k + 1:, k + 1:] -= np.outer(D[k + 1:, k], D[k, k + 1:])
                    Lk = np.tril(D, -1) + np.eye(k1 - k0)
                    Uk = np.triu(D)
                    Fm[k1:, k0:k1] = np.linalg.solve(Uk.T, Fm[k1:, k0:k1].T).T
                    Fm[k0:k1, k1:] = np.linalg.solve(Lk, Fm[k0:k1, k1:])
                    Li, Ui = np.linalg.inv(Lk), np.linalg.inv(Uk)
                    D[:] = np.tril(Li, -1) + np.triu(Ui)
                    upd = Fm[k1:, k0:k1] @ Fm[k0:k1, k1:]
                    mask = np.ones((m - k1, m - k1), dtype=bool)
                    mask[max(0, p - k1):, max(0, p - k1):] = False
                    Fm[k1:, k1:] -= np.where(mask, upd, 0)
                if p < m and p > 0:
                    Fm[p:, p:] -= Fm[p:, :p] @ Fm[:p, p:]
    if stats is not None:
        stats['perturbed'] = nrep
    return arena


def solve(mf, B, arena, x, nb=32):
    """x: (B, n) right-hand sides, solved in place (forward then backward, as the pivot-block
    and panel kernels); ``nb`` as given to :func:`factor`."""
    off, uoff = mf.off, mf.uoff
    widths = _level_widths(mf, nb)
    u = np.zeros(mf.uarena_elems * B, dtype=np.complex128)

    def front(f, b):
        m = int(mf.m[f])
        base = off[f] * B + b * m * m
        return arena[base:base + m * m].reshape(m, m, order='F')
    nlev = mf.lvl_ptr.size - 1
    for L in range(nlev):
        for f in mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]:
            m, p = int(mf.m[f]), int(mf.p[f])
            rows = mf.rows[mf.rows_ptr[f]:mf.rows_ptr[f + 1]]
            for b in range(B):
                Fm = front(f, b).astype(np.complex128)
                w = np.zeros(m, dtype=np.complex128)
                w[:p] = x[b, rows[:p]]
                for slot in range(mf.child_ptr[f], mf.child_ptr[f + 1]):
                    ch = mf.child_ids[slot]
                    inv = mf.inv[mf.inv_ptr[slot]:mf.inv_ptr[slot + 1]]
                    mc = int(mf.m[ch] - mf.p[ch])
                    uc = u[uoff[ch] * B + b * mc: uoff[ch] * B + (b + 1) * mc]
                    sel = inv >= 0
                    w[sel] += uc[inv[sel]]
                nbl = widths[L]
                for k0 in range(0, p, nbl):
                    k1 = min(p, k0 + nbl)
                    w[k0:k1] = (np.tril(Fm[k0:k1, k0:k1], -1) + np.eye(k1 - k0)) @ w[k0:k1]
                    w[k1:p] -= Fm[k1:p, k0:k1] @ w[k0:k1]
                uf = w[p:] - Fm[p:, :p] @ w[:p]
                u[uoff[f] * B + b * (m - p): uoff[f] * B + (b + 1) * (m - p)] = uf
                x[b, rows[:p]] = w[:p]
    for L in range(nlev - 1, -1, -1):
        for f in mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]:
            m, p = int(mf.m[f]), int(mf.p[f])
            rows = mf.rows[mf.rows_ptr[f]:mf.rows_ptr[f + 1]]
            for b in range(B):
                Fm = front(f, b).astype(np.complex128)
                z = x[b, rows].copy()
                z[:p] -= Fm[:p, p:] @ z[p:]
                nbl = widths[L]
                for k0 in reversed(range(0, p, nbl)):
                    k1 = min(p, k0 + nbl)
                    z[k0:k1] = np.triu(Fm[k0:k1, k0:k1]) @ z[k0:k1]
                    z[:k0] -= Fm[:k0, k0:k1] @ z[k0:k1]
                x[b, rows[:p]] = z[:p]
    return x
