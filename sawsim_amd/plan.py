"""Host arrays of the multifrontal numeric plan (input of ``sawfem_mf_create``).

From the symbolic analysis and the active pattern this builds, once per cell:

* per front: order ``m``, pivots ``p``, row list (active ids, pivots first);
* per (front, child): the map parent-row -> child contribution-block index (or -1),
  used by the gather-form extend-add and by the forward solve;
* per A_aa slot: the front that owns it (the one eliminating the earlier of its row and
  column) and its column-major offset inside that front, ordered by level, front and
  offset (consecutive entries of a front land on consecutive addresses);
* levels (leaves first) and the entry ranges of each level.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .symbolic import Symbolic


@dataclass
class MFArrays:
    n: int
    m: np.ndarray
    p: np.ndarray
    rows_ptr: np.ndarray
    rows: np.ndarray
    child_ptr: np.ndarray
    child_ids: np.ndarray
    inv_ptr: np.ndarray
    inv: np.ndarray
    lvl_ptr: np.ndarray
    lvl_fronts: np.ndarray
    ent_front: np.ndarray
    ent_slot: np.ndarray
    ent_loc: np.ndarray
    ent_lvl_ptr: np.ndarray

    @property
    def nf(self):
        return int(self.m.size)

    @property
    def off(self):
        return np.concatenate([[0], np.cumsum(self.m.astype(np.int64) ** 2)])[:-1]

    @property
    def uoff(self):
        return np.concatenate([[0], np.cumsum((self.m - self.p).astype(np.int64))])[:-1]

    @property
    def arena_elems(self):
        return int((self.m.astype(np.int64) ** 2).sum())

    @property
    def uarena_elems(self):
        return int((self.m - self.p).sum())


def build_mf_arrays(sym: Symbolic, aa_indptr: np.ndarray, aa_indices: np.ndarray) -> MFArrays:
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
    return MFArrays(n=n, m=m, p=p, rows_ptr=rows_ptr, rows=rows, child_ptr=child_ptr, child_ids=child_ids,
                    inv_ptr=inv_ptr, inv=inv, lvl_ptr=lvl_ptr, lvl_fronts=lvl_fronts, ent_front=ent_front,
                    ent_slot=ent_slot, ent_loc=ent_loc, ent_lvl_ptr=ent_lvl_ptr)


def panel_width(nb: int, level_pmax: int) -> int:
    """Panel width the factorization uses on a level (narrow panels for small fronts)."""
    return 32 if (nb == 64 and level_pmax <= 128) else nb


def kernel_work(mf: MFArrays, nb: int = 32) -> dict:
    """Algorithmic work per frequency, counted from the fronts, for panel width ``nb``.

    * ``factor_flops``: 8·Σ_f Σ_{k<p} (m-k-1)² (complex LU, 8 real flops per complex
      multiply-add);
    * ``lupd_flops``: L-shaped panel updates 8·kb·((m-k1)² - (m-p)²) per panel (all panels
      but the last of a front), with the level's panel width;
    * ``schur_flops``: Schur complement 8·p·(m-p)²;
    * ``gemm_flops``: their sum (the MFMA update kernels);
    * ``solve_flops``: 8·(nnz(L)+nnz(U)) for one forward + backward substitution;
    * ``lu_entries``: stored L/U entries (pivot triangle + both panels), the data one solve
      streams; ``arena_entries``: Σ m²;
    * ``schur_bytes``: the Schur complement's algorithmic complex64 bytes, 8·(2·p·q + 2·q²) with
      q = m - p (L21 and U12 read once, the contribution block read and written once).
    """
    m = mf.m.astype(np.float64)
    p = mf.p.astype(np.float64)
    ff = 0.0
    lf = 0.0
    sf = 0.0
    for L in range(mf.lvl_ptr.size - 1):
        fronts = mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]
        nbl = panel_width(nb, int(mf.p[fronts].max()))
        for f in fronts:
            mi, pi = int(mf.m[f]), int(mf.p[f])
            k = np.arange(pi, dtype=np.float64)
            ff += float(((mi - k - 1) ** 2).sum())
            for k0 in range(0, pi, nbl):
                k1 = min(pi, k0 + nbl)
                if k1 < pi:
                    lf += (k1 - k0) * float((mi - k1) ** 2 - (mi - pi) ** 2)
            sf += float(pi) * float((mi - pi) ** 2)
    nnz_lu = float((p * p + 2 * p * (m - p)).sum())
    q = m - p
    return dict(factor_flops=8 * ff, lupd_flops=8 * lf, schur_flops=8 * sf, gemm_flops=8 * (lf + sf),
                solve_flops=8 * nnz_lu, lu_entries=nnz_lu, arena_entries=float((m * m).sum()),
                schur_bytes=float((8 * (2 * p * q + 2 * q * q))[(q > 0) & (p > 0)].sum()))
