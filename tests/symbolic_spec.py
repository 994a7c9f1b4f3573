"""Pure-NumPy statement of the host symbolic analysis and the plan arrays.

The library computes both in C++ (``sawfem_analyse``, ``csrc/analyse.cpp``); this is the
specification those arrays are checked against, bit for bit, on the CPU
(``tests/test_analyse.py``).  Nested-dissection ordering on supervariable coordinates,
periodic in x; fronts in post order; rows of each front in elimination order; every A_aa entry's
owner front and column-major offset, sorted by level, front and offset.
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


def analyse(aa_indptr, aa_indices, node_of, comp_of, node_xy, period=None, leaf_weight=96):
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
    sym.stats = dict(root=root)
    return sym


def build_mf_arrays(sym: Symbolic, aa_indptr: np.ndarray, aa_indices: np.ndarray) -> dict:
    F = sym.fronts
    nf = len(F)
    n = sym.n
    m = np.array([f.m for f in F], dtype=np.int32)
    p = np.array([f.p for f in F], dtype=np.int32)
    rows_ptr = np.concatenate([[0], np.cumsum(m.astype(np.int64))])
    rows = np.concatenate([f.rows for f in F]).astype(np.int32)
    child_ptr = np.concatenate([[0], np.cumsum([len(f.children) for f in F])]).astype(np.int32)
    child_ids = np.array([c for f in F for c in f.children], dtype=np.int32)
    # parent-row -> child CB index maps
    scratch = np.full(n, -1, dtype=np.int32)
    inv_parts = []
    for f in F:
        for c in f.children:
            g = F[c]
            cb = g.rows[g.p:]
            scratch[cb] = np.arange(cb.size, dtype=np.int32)
            inv = scratch[f.rows].copy()
            if (inv >= 0).sum() != cb.size:
                raise AssertionError('child contribution rows missing from parent front')
            inv_parts.append(inv)
            scratch[cb] = -1
    inv_ptr = np.concatenate([[0], np.cumsum([x.size for x in inv_parts])]).astype(np.int64)
    inv = np.concatenate(inv_parts).astype(np.int32) if inv_parts else np.zeros(0, dtype=np.int32)
    # levels
    lvl_fronts = np.concatenate(sym.levels).astype(np.int32)
    lvl_ptr = np.concatenate([[0], np.cumsum([lv.size for lv in sym.levels])]).astype(np.int32)
    level_of = np.empty(nf, dtype=np.int64)
    for L, lv in enumerate(sym.levels):
        level_of[lv] = L
    # entry ownership
    owner = np.empty(n, dtype=np.int64)
    for fid, f in enumerate(F):
        owner[f.rows[:f.p]] = fid
    rank = np.empty(n, dtype=np.int64)
    rank[sym.perm] = np.arange(n)
    r = np.repeat(np.arange(n, dtype=np.int64), np.diff(aa_indptr))
    c = aa_indices.astype(np.int64)
    own = np.where(rank[r] <= rank[c], owner[r], owner[c])
    # local index lookup table keyed by front * n + active id
    fk = np.repeat(np.arange(nf, dtype=np.int64), m.astype(np.int64)) * n + rows.astype(np.int64)
    lk = np.concatenate([np.arange(x, dtype=np.int64) for x in m])
    o = np.argsort(fk)
    fk, lk = fk[o], lk[o]

    def local(front, ids):
        key = front * n + ids
        pos = np.searchsorted(fk, key)
        if not np.all(fk[np.minimum(pos, fk.size - 1)] == key):
            raise AssertionError('A_aa entry outside its owner front')
        return lk[pos]
    li, lj = local(own, r), local(own, c)
    loc = li + lj * m[own].astype(np.int64)
    order = np.lexsort((loc, own, level_of[own]))      # by level, front, column-major offset
    ent_front = own[order].astype(np.int32)
    ent_slot = np.arange(r.size, dtype=np.int64)[order]
    ent_loc = loc[order].astype(np.int32)
    ent_lvl_ptr = np.searchsorted(level_of[ent_front], np.arange(len(sym.levels) + 1)).astype(np.int64)
    return dict(n=n, m=m, p=p, rows_ptr=rows_ptr, rows=rows, child_ptr=child_ptr, child_ids=child_ids,
                inv_ptr=inv_ptr, inv=inv, lvl_ptr=lvl_ptr, lvl_fronts=lvl_fronts, ent_front=ent_front,
                ent_slot=ent_slot, ent_loc=ent_loc, ent_lvl_ptr=ent_lvl_ptr)


def _gather_lists(targets, n_out):
    order = np.argsort(targets, kind='stable')
    counts = np.bincount(targets, minlength=n_out)
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.int64), order.astype(np.int64)


def bloch_arrays(kl, rep, pos, n, is_z, is_f) -> dict:
    """The Bloch layout's pattern and gather lists over the K slots (``sweep.py:293-309``):
    A_aa pattern of the merged active rows/columns, per A_aa entry its K slots; per active row the
    K slots of the electrode column (Km@s); the charge weights on the unmerged K."""
    rows, cols = kl.slot_rows_cols()
    mr, mc = rep[rows], rep[cols]
    pr, pc = pos[mr], pos[mc]
    sel = np.flatnonzero((pr >= 0) & (pc >= 0))
    keys = pr[sel] * n + pc[sel]
    ukeys, inv = np.unique(keys, return_inverse=True)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, ukeys // n + 1, 1)
    out = dict(aa_indptr=np.cumsum(indptr), aa_indices=(ukeys % n).astype(np.int64))
    ptr, order = _gather_lists(inv, ukeys.size)
    out.update(aa_ptr=ptr, aa_idx=sel[order])
    sel = np.flatnonzero((pr >= 0) & is_z[mc])
    ptr, order = _gather_lists(pr[sel], n)
    out.update(rhs_ptr=ptr, rhs_idx=sel[order])
    zr = is_z[rows]
    sel = np.flatnonzero(zr & (pc >= 0) & ~is_z[cols] & ~is_f[cols])
    selz = np.flatnonzero(zr & is_z[cols])
    tgt = np.concatenate([pc[sel], np.full(selz.size, n)])
    ptr, order = _gather_lists(tgt, n + 1)
    out.update(kts_ptr=ptr, kts_idx=np.concatenate([sel, selz])[order])
    return out
