"""Host arrays of the multifrontal numeric plan (input of ``sawfem_mf_create``).

``sawfem_analyse`` (``csrc/analyse.cpp``) builds them with the ordering, once per cell:

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
