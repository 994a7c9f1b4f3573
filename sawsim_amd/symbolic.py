"""Host symbolic analysis for the multifrontal LU of the Bloch-reduced system.

Done once per cell (all frequencies and, for a design batch, all cells of the same
mesh share it), replacing the analysis phase of PARDISO / SuperLU
(``sweep.py:326-333, 349-353`` call them once per frequency).

1. Supervariables: the active unknowns of one merged mesh node share their sparsity
   pattern, so the ordering works on the merged-node graph.
2. Ordering: geometric nested dissection on node coordinates, periodic in x.  A
   sub-domain is split at the weighted median of y, or of x (for a sub-domain that
   wraps around the periodic cell, two cuts half a period apart); the vertex
   separator is the boundary layer of one side, and the cut that gives the lighter
   separator wins.  Each separator (and each leaf sub-domain) becomes one front.
3. Fronts: the dissection tree is the assembly tree.  A front's row set is its
   pivots, the pattern neighbours of its pivots and its children's contribution
   rows, restricted to unknowns eliminated later.

Everything is expanded to scalar unknowns at the end; front rows are listed as
positions in the reference active ordering (``sweep.py:298``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Front:
    pivots_sv: np.ndarray           # supervariables eliminated here
    children: list
    parent: int = -1
    rows_sv: np.ndarray = None      # pivots then contribution supervariables (elimination order)
    rows: np.ndarray = None         # active-unknown ids (reference order), pivots first
    p: int = 0                      # pivot count (scalar unknowns)
    m: int = 0                      # front order
    level: int = 0                  # 0 = leaf level


@dataclass
class Symbolic:
    fronts: list
    order_sv: np.ndarray            # elimination order of supervariables
    perm: np.ndarray                # elimination order of active unknowns
    levels: list                    # list of front-id arrays, leaves first
    n: int
    stats: dict = field(default_factory=dict)


def _sv_graph(aa_indptr, aa_indices, node_of):
    """Supervariable adjacency (CSR, no self loops) from the active pattern."""
    uniq, sv_of = np.unique(node_of, return_inverse=True)
    n = node_of.size
    rows = np.repeat(np.arange(n), np.diff(aa_indptr))
    a, b = sv_of[rows], sv_of[aa_indices]
    nsv = uniq.size
    keys = np.unique(a[a != b] * nsv + b[a != b])
    ptr = np.zeros(nsv + 1, dtype=np.int64)
    np.add.at(ptr, keys // nsv + 1, 1)
    return uniq, sv_of, np.cumsum(ptr), (keys % nsv).astype(np.int64)


def _neighbours(ptr, idx, verts):
    cnt = ptr[verts + 1] - ptr[verts]
    src = np.repeat(verts, cnt)
    offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    return src, idx[np.repeat(ptr[verts], cnt) + offs]


class _Dissector:
    def __init__(self, ptr, idx, x, y, w, period, leaf_weight):
        self.ptr, self.idx, self.x, self.y, self.w = ptr, idx, x, y, w
        self.period = period
        self.leaf = leaf_weight
        self.label = np.full(x.size, -1, dtype=np.int64)
        self.tag = 0

    def _separator(self, verts, side_a):
        """Lighter of the two one-sided boundary layers between parts A and B."""
        self.tag += 2
        ta, tb = self.tag, self.tag + 1
        self.label[verts[side_a]] = ta
        self.label[verts[~side_a]] = tb
        src, dst = _neighbours(self.ptr, self.idx, verts)
        la, lb = self.label[src], self.label[dst]
        sa = np.unique(src[(la == ta) & (lb == tb)])
        sb = np.unique(src[(la == tb) & (lb == ta)])
        if self.w[sa].sum() <= self.w[sb].sum():
            return sa
        return sb

    def _weighted_median(self, key, verts):
        o = np.argsort(key, kind='stable')
        cw = np.cumsum(self.w[verts][o])
        k = int(np.searchsorted(cw, cw[-1] / 2))
        return key[o[min(k, o.size - 1)]]

    def candidates(self, verts):
        x, y = self.x[verts], self.y[verts]
        out = []
        ym = self._weighted_median(y, verts)
        out.append(y < ym)
        if self.period is not None:
            P = self.period
            xs = np.sort(np.unique(np.round(x / P * 1e9)) / 1e9 * P)
            gaps = np.diff(np.concatenate([xs, [xs[0] + P]]))
            g = int(np.argmax(gaps))
            wraps = gaps[g] < 0.26 * P if xs.size > 1 else False
            if wraps:
                xu = np.mod(x, P)
                out.append(xu < P / 2 - 1e-15 * P)
            else:
                start = xs[(g + 1) % xs.size]
                xu = np.mod(x - start + 1e-12 * P, P)
                xm = self._weighted_median(xu, verts)
                out.append(xu < xm)
        else:
            xm = self._weighted_median(x, verts)
            out.append(x < xm)
        return out

    def run(self, verts):
        """Returns a nested tuple tree: (pivots, [children]) (leaf: children == [])."""
        if self.w[verts].sum() <= self.leaf or verts.size <= 2:
            return (verts, [])
        best = None
        for side_a in self.candidates(verts):
            if side_a.all() or (~side_a).all():
                continue
            sep = self._separator(verts, side_a)
            rest = np.setdiff1d(verts, sep)
            ina = np.isin(rest, verts[side_a])
            pa, pb = rest[ina], rest[~ina]
            bal = min(self.w[pa].sum(), self.w[pb].sum()) / max(1, self.w[verts].sum())
            score = self.w[sep].sum() / max(bal, 1e-3)
            if best is None or score < best[0]:
                best = (score, sep, pa, pb)
        if best is None:
            return (verts, [])
        _, sep, pa, pb = best
        kids = [self.run(part) for part in (pa, pb) if part.size]
        if sep.size == 0:
            return (np.zeros(0, dtype=np.int64), kids)
        return (sep, kids)


def analyse(aa_indptr, aa_indices, node_of, comp_of, node_xy, period=None, leaf_weight=96,
            amalgamate=0):
    """Nested-dissection multifrontal analysis of the active system.

    Parameters
    ----------
    aa_indptr, aa_indices : active pattern (reference active ordering).
    node_of, comp_of : merged mesh node and component of each active unknown.
    node_xy : (N, 2) node coordinates.
    period : cell width for Bloch-periodic x, or None.
    leaf_weight : maximum scalar unknowns in a leaf sub-domain.
    """
    n = node_of.size
    uniq, sv_of, ptr, idx = _sv_graph(aa_indptr, aa_indices, node_of)
    nsv = uniq.size
    w = np.bincount(sv_of, minlength=nsv)
    x, y = node_xy[uniq, 0], node_xy[uniq, 1]
    dis = _Dissector(ptr, idx, x, y, w, period, leaf_weight)
    tree = dis.run(np.arange(nsv))

    fronts = []

    def build(node):
        piv, kids = node
        cids = [build(k) for k in kids]
        cids = [c for c in cids if c is not None]
        if piv.size == 0:
            # empty separator: hoist children (disconnected parts) under a pivot-less node
            if not cids:
                return None
            if len(cids) == 1:
                return cids[0]
        f = Front(pivots_sv=piv, children=cids)
        fronts.append(f)
        fid = len(fronts) - 1
        for c in cids:
            fronts[c].parent = fid
        return fid
    root = build(tree)
    # elimination order: fronts are appended in post-order
    order_sv = np.concatenate([f.pivots_sv for f in fronts])
    pos_sv = np.empty(nsv, dtype=np.int64)
    pos_sv[order_sv] = np.arange(nsv)
    start = np.cumsum([0] + [f.pivots_sv.size for f in fronts])
    # dofs of each supervariable in component order
    dof_order = np.lexsort((comp_of, pos_sv[sv_of]))     # active unknowns in elimination order
    sv_dof_ptr = np.concatenate([[0], np.cumsum(np.bincount(sv_of[dof_order], minlength=nsv)[order_sv])])
    # row structure, bottom-up
    for fid, f in enumerate(fronts):
        s0 = start[fid]
        _, nb = _neighbours(ptr, idx, f.pivots_sv)
        parts = [f.pivots_sv, nb] + [fronts[c].rows_sv[fronts[c].pivots_sv.size:] for c in f.children]
        cand = np.unique(np.concatenate(parts))
        cand = cand[pos_sv[cand] >= s0]
        f.rows_sv = cand[np.argsort(pos_sv[cand], kind='stable')]
    # expand to scalar unknowns
    for f in fronts:
        p_sv = pos_sv[f.rows_sv]
        lo, hi = sv_dof_ptr[p_sv], sv_dof_ptr[p_sv + 1]
        cnt = hi - lo
        offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        f.rows = dof_order[np.repeat(lo, cnt) + offs]
        f.m = int(f.rows.size)
        f.p = int(cnt[:f.pivots_sv.size].sum())
    for fid, f in enumerate(fronts):
        f.level = 0 if not f.children else 1 + max(fronts[c].level for c in f.children)
    nlev = max(f.level for f in fronts) + 1
    levels = [np.array([i for i, f in enumerate(fronts) if f.level == L], dtype=np.int64)
              for L in range(nlev)]
    perm = np.concatenate([f.rows[:f.p] for f in fronts])
    assert perm.size == n and np.unique(perm).size == n
    sym = Symbolic(fronts=fronts, order_sv=order_sv, perm=perm, levels=levels, n=n)
    sym.stats = front_stats(sym)
    sym.stats['root'] = root
    return sym


def front_stats(sym: Symbolic) -> dict:
    """Algorithmic work of the numeric phase, counted from the fronts.

    ``flops_factor`` = 8·Σ_fronts Σ_{k<p} (m-k-1)² real flops of complex LU (one
    complex multiply-add = 8 real flops); ``nnz_lu`` = stored L and U entries.
    """
    ff, nnz, cb = 0, 0, 0
    for f in sym.fronts:
        m, p = f.m, f.p
        k = np.arange(p, dtype=np.float64)
        ff += float(((m - k - 1) ** 2).sum())
        nnz += p * p + 2 * p * (m - p)
        cb += (m - p) ** 2
    ms = np.array([f.m for f in sym.fronts])
    ps = np.array([f.p for f in sym.fronts])
    return dict(n_fronts=len(sym.fronts), n_levels=len(sym.levels), flops_factor=8.0 * ff,
                nnz_lu=int(nnz), cb_entries=int(cb), max_m=int(ms.max()), max_p=int(ps.max()),
                sum_m2=int((ms.astype(np.int64) ** 2).sum()))
