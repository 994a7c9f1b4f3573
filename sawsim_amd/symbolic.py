"""Host symbolic analysis for the multifrontal LU of the Bloch-reduced system.

Done once per cell (all frequencies and, for a design batch, all cells of the same mesh share
it), replacing the analysis phase of PARDISO / SuperLU (``sweep.py:326-333, 349-353`` call them
once per frequency).  The work is native: ``sawfem_analyse`` (``csrc/analyse.cpp``) runs

1. supervariables: the active unknowns of one merged mesh node share their sparsity pattern,
   so the ordering works on the merged-node graph;
2. ordering: geometric nested dissection on node coordinates, periodic in x.  A sub-domain is
   split at the weighted median of y, or of x (for a sub-domain that wraps around the periodic
   cell, two cuts half a period apart); the vertex separator is the boundary layer of one side,
   and the cut that gives the lighter separator per balance wins.  Each separator (and each
   leaf sub-domain) becomes one front;
3. fronts: the dissection tree is the assembly tree.  A front's row set is its pivots, the
   pattern neighbours of its pivots and its children's contribution rows, restricted to
   unknowns eliminated later;
4. the numeric plan's arrays (:class:`plan.MFArrays`).

Front rows are positions in the reference active ordering (``sweep.py:298``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from functools import cached_property

import numpy as np

from . import _lib
from .plan import MFArrays

# array ids of sawfem_host_copy (include/sawfem.h) and their element types
_ARRAYS = dict(m=(0, np.int32), p=(1, np.int32), rows_ptr=(2, np.int64), rows=(3, np.int32),
               child_ptr=(4, np.int32), child_ids=(5, np.int32), parent=(6, np.int32), level=(7, np.int32),
               inv_ptr=(8, np.int64), inv=(9, np.int32), lvl_ptr=(10, np.int32), lvl_fronts=(11, np.int32),
               ent_front=(12, np.int32), ent_slot=(13, np.int64), ent_loc=(14, np.int32),
               ent_lvl_ptr=(15, np.int64), perm=(16, np.int64), piv_sv_ptr=(17, np.int64), piv_sv=(18, np.int64),
               rows_sv_ptr=(19, np.int64), rows_sv=(20, np.int64), sv_ids=(21, np.int64))


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


class Symbolic:
    """Result of :func:`analyse`: the ordering, the fronts and the plan arrays (``mf``)."""

    def __init__(self, a: dict, n: int):
        self.arrays = a
        self.n = n
        self.perm = a['perm']
        self.order_sv = a['sv_ids'][a['piv_sv']]         # supervariables as sv_of values
        self.mf = MFArrays(n=n, **{k: a[k] for k in ('m', 'p', 'rows_ptr', 'rows', 'child_ptr', 'child_ids',
                                                      'inv_ptr', 'inv', 'lvl_ptr', 'lvl_fronts', 'ent_front',
                                                      'ent_slot', 'ent_loc', 'ent_lvl_ptr')})
        lp = a['lvl_ptr']
        self.levels = [a['lvl_fronts'][lp[L]:lp[L + 1]].astype(np.int64) for L in range(lp.size - 1)]
        self.stats = front_stats(self.mf)
        roots = np.flatnonzero(a['parent'] < 0)
        self.stats['root'] = int(roots[-1])

    @cached_property
    def fronts(self):
        """Per-front view (built on first use)."""
        a = self.arrays
        out = []
        for f in range(a['m'].size):
            out.append(Front(pivots_sv=a['piv_sv'][a['piv_sv_ptr'][f]:a['piv_sv_ptr'][f + 1]],
                             children=a['child_ids'][a['child_ptr'][f]:a['child_ptr'][f + 1]].tolist(),
                             parent=int(a['parent'][f]),
                             rows_sv=a['rows_sv'][a['rows_sv_ptr'][f]:a['rows_sv_ptr'][f + 1]],
                             rows=a['rows'][a['rows_ptr'][f]:a['rows_ptr'][f + 1]].astype(np.int64),
                             p=int(a['p'][f]), m=int(a['m'][f]), level=int(a['level'][f])))
        return out


def analyse(aa_indptr, aa_indices, node_of, comp_of, node_xy, period=None, leaf_weight=96) -> Symbolic:
    """Nested-dissection multifrontal analysis of the active system (``sawfem_analyse``).

    Parameters
    ----------
    aa_indptr, aa_indices : active pattern (reference active ordering).
    node_of, comp_of : merged mesh node (supervariable id into ``node_xy``) and component of
        each active unknown.
    node_xy : (N, 2) node coordinates.
    period : cell width for Bloch-periodic x, or None.
    leaf_weight : maximum scalar unknowns in a leaf sub-domain.
    """
    n = int(np.asarray(node_of).size)
    i64 = lambda v: np.ascontiguousarray(v, dtype=np.int64)
    ip, ix, sv, cp = i64(aa_indptr), i64(aa_indices), i64(node_of), i64(comp_of)
    xy = np.ascontiguousarray(np.asarray(node_xy, dtype=np.float64)[:, :2])
    if ip.size != n + 1 or ix.size != ip[-1] or cp.size != n:
        raise ValueError('pattern, supervariable and component arrays do not match')
    h = C.c_void_p()
    _lib.call('sawfem_analyse', C.byref(h), n, ip.ctypes.data, ix.ctypes.data, sv.ctypes.data, cp.ctypes.data,
              xy.shape[0], xy.ctypes.data, float(period or 0.0), int(leaf_weight))
    return Symbolic(_lib.host_arrays(h, _ARRAYS), n)


def front_stats(mf: MFArrays) -> dict:
    """Algorithmic work of the numeric phase, counted from the fronts.

    ``flops_factor`` = 8·Σ_fronts Σ_{k<p} (m-k-1)² real flops of complex LU (one complex
    multiply-add = 8 real flops), in closed form Σ_{j=m-p}^{m-1} j²; ``nnz_lu`` = stored L and U
    entries.
    """
    m = mf.m.astype(np.int64)
    p = mf.p.astype(np.int64)
    sq = lambda k: (k - 1) * k * (2 * k - 1) // 6                # Σ_{j<k} j²
    ff = (sq(m) - sq(m - p)).astype(np.float64).sum()
    return dict(n_fronts=int(m.size), n_levels=int(mf.lvl_ptr.size - 1), flops_factor=8.0 * ff,
                nnz_lu=int((p * p + 2 * p * (m - p)).sum()), cb_entries=int(((m - p) ** 2).sum()),
                max_m=int(m.max()), max_p=int(p.max()), sum_m2=int((m * m).sum()))
