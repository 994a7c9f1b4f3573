"""Oracle: periodic boundary sets, merge transform and the SuperLU frequency solve
(test infrastructure only).

Restates ``saw2d/sweep.py``:

* ``bc_periodic``       -> :41-72 (lines 17 +, 18 -, 19 left, 20 right, 21 bottom)
* ``merge_transform``   -> :97-144 (union-find with path halving, first master wins)
* ``real_block``        -> :256-262
* ``FrequencySolver``   -> :275-364 (T^T K T, active set, kts = K^T s, damping :316-323)
* ``sweep_charges``     -> ``frequency_tasks.py:12-55`` fork pool, charges in order
"""
from __future__ import annotations

import multiprocessing as mp
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def nodes_with_tag(lines, tag):
    return np.unique(lines[lines[:, 3] == tag][:, :3].astype(np.int64) - 1)


def bc_periodic(pos_m, lines, N, ndisp=3, tag_bottom=21, tag_left=19, tag_right=20,
                tag_zheng=17, tag_fu=18):
    """(fix, zheng, fu, merge_groups) for a Bloch-periodic SP cell (sweep.py:41-72)."""
    bottom = nodes_with_tag(lines, tag_bottom)
    left = nodes_with_tag(lines, tag_left)
    right = nodes_with_tag(lines, tag_right)
    zheng = nodes_with_tag(lines, tag_zheng)
    fu = nodes_with_tag(lines, tag_fu)
    cL = np.intersect1d(bottom, left)
    cR = np.intersect1d(bottom, right)
    L = np.setdiff1d(left, cL)
    R = np.setdiff1d(right, cR)
    L = L[np.argsort(pos_m[L, 1])]
    R = R[np.argsort(pos_m[R, 1])]
    if L.size != R.size:
        raise RuntimeError('left/right periodic node counts differ')
    masters = np.concatenate([L + a * N for a in range(ndisp)] + [L + ndisp * N, cL + ndisp * N])
    slaves = np.concatenate([R + a * N for a in range(ndisp)] + [R + ndisp * N, cR + ndisp * N])
    groups = [np.array([m, s]) for m, s in zip(masters, slaves)]
    fix = np.concatenate([bottom + a * N for a in range(ndisp)])
    return fix, zheng + ndisp * N, fu + ndisp * N, groups


def merge_transform(gdof, groups):
    """T (gdof x gdof, 0/1), slaves, masters such that A_merged = Tᵀ A T (sweep.py:97-144)."""
    members = []
    seen = set()
    links = []
    for g in groups:
        g = [int(x) for x in np.asarray(g, dtype=int).ravel()]
        if not g:
            continue
        for x in g:
            if x not in seen:
                seen.add(x)
                members.append(x)
        links += [(s, g[0]) for s in g[1:]]
    parent = {x: x for x in members}

    def root(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for s, m in links:
        rs, rm = root(s), root(m)
        if rs != rm:
            parent[rs] = rm
    # the reference iterates a Python set of ints; order of (slave, master) pairs is
    # irrelevant to T, and we sort by slave for a canonical listing
    pairs = sorted((x, root(x)) for x in members if root(x) != x)
    slaves = np.array([p[0] for p in pairs], dtype=np.int64)
    masters = np.array([p[1] for p in pairs], dtype=np.int64)
    is_slave = np.zeros(gdof, dtype=bool)
    is_slave[slaves] = True
    keep = np.flatnonzero(~is_slave)
    rows = np.concatenate([keep, slaves])
    cols = np.concatenate([keep, masters])
    T = sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(gdof, gdof)).tocsr()
    return T, slaves, masters


def real_block(A):
    """[[Ar, -Ai], [Ai, Ar]] (sweep.py:256-262)."""
    out = sp.bmat([[A.real, -A.imag], [A.imag, A.real]], format='csr')
    out.sort_indices()
    return out


class FrequencySolver:
    """Per-frequency solve of (K + iωβD − ω²M) with Bloch merge and SuperLU (sweep.py:275-364)."""

    def __init__(self, K, M, fix, zheng, fu, groups):
        self.K = K.tocsr()
        self.gdof = gdof = K.shape[0]
        self.zheng = np.asarray(zheng, dtype=np.int64)
        self.fu = np.asarray(fu, dtype=np.int64)
        T, self.slaves, self.masters = merge_transform(gdof, groups)
        self.T = T
        Km = (T.T @ self.K @ T).tocsr()
        Mm = (T.T @ M.tocsr() @ T).tocsr()
        prescribed = np.unique(np.concatenate([np.asarray(fix, dtype=np.int64), self.slaves,
                                               self.zheng, self.fu]))
        self.active = np.setdiff1d(np.arange(gdof), prescribed)
        self.n_active = self.active.size
        s = np.zeros(gdof)
        s[self.zheng] = 1.0
        self.Km_s = np.asarray(Km @ s).ravel()
        self.Mm_s = np.asarray(Mm @ s).ravel()
        self.kts = np.asarray(self.K.T @ s).ravel()
        self.Km_aa = Km[self.active][:, self.active].tocsr()
        self.Mm_aa = Mm[self.active][:, self.active].tocsr()
        self.beta = 0.0

    def set_damping(self, D, beta):
        Dm = (self.T.T @ D.tocsr() @ self.T).tocsr()
        s = np.zeros(self.gdof)
        s[self.zheng] = 1.0
        self.Dm_s = np.asarray(Dm @ s).ravel()
        self.Dm_aa = Dm[self.active][:, self.active].tocsr()
        self.beta = float(beta)

    def system(self, omega, V=1.0):
        """(A_aa, rhs) for one angular frequency (sweep.py:337-346)."""
        w2 = omega ** 2
        Ks, Ka = self.Km_s, self.Km_aa
        if self.beta:
            Ks = Ks + 1j * omega * self.beta * self.Dm_s
            Ka = Ka + 1j * omega * self.beta * self.Dm_aa
        rhs = -(Ks - w2 * self.Mm_s)[self.active] * V
        return (Ka - w2 * self.Mm_aa).tocsr(), rhs

    def expand(self, x, V=1.0):
        disp = np.zeros(self.gdof, dtype=complex)
        disp[self.active] = x
        disp[self.slaves] = disp[self.masters]
        disp[self.zheng] = V
        disp[self.fu] = 0.0
        return disp

    def solve(self, omega, V=1.0):
        """(disp, Q, |iωQ/V|) via the real-block system and SuperLU (sweep.py:335-364)."""
        A, rhs = self.system(omega, V)
        n = self.n_active
        z = spla.spsolve(real_block(A).tocsc(), np.concatenate([rhs.real, rhs.imag]))
        disp = self.expand(z[:n] + 1j * z[n:], V)
        Q = self.kts @ disp
        return disp, Q, abs(1j * omega * Q / V)

    def solve_complex(self, omega, V=1.0):
        """Same system solved directly in complex arithmetic (SuperLU), for cross-checks."""
        A, rhs = self.system(omega, V)
        disp = self.expand(spla.spsolve(A.tocsc(), rhs), V)
        Q = self.kts @ disp
        return disp, Q, abs(1j * omega * Q / V)


_POOL_ENGINE = None


def _pool_charge(args):
    f, V, cplx = args
    solve = _POOL_ENGINE.solve_complex if cplx else _POOL_ENGINE.solve
    return complex(solve(2 * np.pi * f, V)[1])


def sweep_charges(engine, freqs, V=1.0, workers=None, complex_form=False):
    """Charges in frequency order over a fork pool (frequency_tasks.py:12-55).

    ``complex_form`` solves the complex system directly instead of the real-block form."""
    global _POOL_ENGINE
    freqs = np.asarray(freqs, dtype=float)
    if workers is None:
        workers = 8 if freqs.size > 41 else 1
    workers = max(1, min(int(workers), freqs.size, os.cpu_count() or 1))
    if workers == 1:
        solve = engine.solve_complex if complex_form else engine.solve
        return np.array([solve(2 * np.pi * f, V)[1] for f in freqs], dtype=complex)
    _POOL_ENGINE = engine
    with mp.get_context('fork').Pool(workers) as pool:
        q = pool.map(_pool_charge, [(f, V, complex_form) for f in freqs], chunksize=1)
    return np.asarray(q, dtype=complex)
