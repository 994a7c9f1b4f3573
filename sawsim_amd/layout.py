"""Host-side integer layout of the assembled and Bloch-reduced systems.

Everything here is built once per cell on the host and uploaded as index arrays;
the floating-point work that consumes it runs in the HIP kernels.

DOF numbering follows the reference (``piezo_fem.py:11, 319-325``): component-major,
``dof = comp * N + node``; ``nd = 3`` for the plane-strain ME0 system ``[ux|uy|phi]``
(``piezo_fem.py:402-405``) and ``nd = 4`` for ME1 ``[ux|uy|uz|phi]``.

Patterns:

* ``K`` (and ``M``, ``D``) share one *structural* CSR pattern: every component pair of
  every node pair that shares an element.  The reference builds K and M by COO -> CSR
  (``piezo_fem.py:352-365``), which keeps explicit zeros, so its K pattern is this
  structural one; its M keeps only the diagonal displacement blocks, which here are
  slots holding exact zeros.
* ``A_aa`` is ``(Tᵀ K T)[active, active]`` (``sweep.py:293-309``) on the structural
  pattern of the merged active unknowns.  SciPy's SpGEMM drops exact-zero sums, so the
  reference pattern is a value-dependent subset of this one; the extra slots hold 0.

Gather lists (``ptr``, ``idx``) drive a deterministic, atomic-free reduction:
``out[j] = sum(in[idx[k]] for k in ptr[j]:ptr[j+1])`` in list order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


def _csr_from_sorted_keys(keys, nrows, ncols):
    rows = keys // ncols
    cols = keys % ncols
    indptr = np.zeros(nrows + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    return np.cumsum(indptr), cols.astype(np.int64)


def _gather_lists(targets, n_out):
    """Group source positions by target id (stable): ptr (n_out+1), idx (sources)."""
    order = np.argsort(targets, kind='stable')
    counts = np.bincount(targets, minlength=n_out)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return ptr, order.astype(np.int64)


@dataclass
class GatherList:
    ptr: np.ndarray
    idx: np.ndarray

    @property
    def n_out(self):
        return self.ptr.size - 1


class KLayout:
    """Structural CSR pattern of K/M/D and the element -> slot gather (K1/K2 kernels).

    Element matrices are written by the element kernels as ``Ke[e, A, B, i, j]`` (complex,
    component blocks of npe x npe node entries, npe = 9 for Q9 and 27 for Hex27) and
    ``Me[e, i, j]``.  The global matrices are then gathered per *node pair* ``k`` (an entry
    of the node adjacency) from the list of ``(e, i, j)`` codes ``e*npe² + i*npe + j`` of the
    elements that couple that node pair.

    Storage is block-major: slot = A*nd*nnz_node + nd*node_ptr[i] + B*deg[i] + (k - node_ptr[i]).
    ``numbering`` maps (component A, node) to a global dof: ``'component'`` = A*N + node
    (the 2D reference, ``piezo_fem.py:11``), ``'interleaved'`` = 3*node + A for the three
    displacements and 3N + node for the potential (the Hex27 reference, ``saw3d/assembly.py:4-5``).
    """

    def __init__(self, quads: np.ndarray, N: int, nd: int, numbering: str = 'component'):
        quads = np.asarray(quads, dtype=np.int64)
        ne, npe = quads.shape
        if numbering not in ('component', 'interleaved') or (numbering == 'interleaved' and nd != 4):
            raise ValueError(f'unsupported numbering {numbering!r} for nd={nd}')
        self.N, self.nd, self.ne, self.npe = int(N), int(nd), int(ne), int(npe)
        self.numbering = numbering
        ri = np.repeat(quads, npe, axis=1)               # (ne,npe²) row node of local (i,j)
        ci = np.tile(quads, (1, npe))                     # (ne,npe²) col node
        keys = (ri * N + ci).ravel()
        ukeys, inv = np.unique(keys, return_inverse=True)
        self.node_ptr, self.node_idx = _csr_from_sorted_keys(ukeys, N, N)
        self.nnz_node = int(ukeys.size)
        self.pair_row = (ukeys // N).astype(np.int64)     # row node of each node pair
        self.pair_col = (ukeys % N).astype(np.int64)
        self.elem_pair = inv.reshape(ne, npe * npe).astype(np.int64)   # node-pair slot per element entry
        ptr, idx = _gather_lists(self.elem_pair.ravel(), self.nnz_node)
        self.pair_gather = GatherList(ptr, idx)           # codes e*npe² + i*npe + j
        deg = np.diff(self.node_ptr)
        self.deg = deg
        self.nnz = nd * nd * self.nnz_node
        self.gdof = nd * N

    def dof(self, a, node):
        """Global dof of component ``a`` at ``node``."""
        node = np.asarray(node, dtype=np.int64)
        if self.numbering == 'component':
            return a * self.N + node
        return 3 * node + a if a < 3 else 3 * self.N + node

    def node_comp(self, dofs):
        """(node, component) of global dofs."""
        dofs = np.asarray(dofs, dtype=np.int64)
        if self.numbering == 'component':
            return dofs % self.N, dofs // self.N
        u = dofs < 3 * self.N
        return np.where(u, dofs // 3, dofs - 3 * self.N), np.where(u, dofs % 3, 3)

    def slot(self, a, b, pair):
        """K slot of component block (a, b) at node-pair slot(s) ``pair``."""
        i = self.pair_row[pair]
        return (a * self.nd * self.nnz_node + self.nd * self.node_ptr[i] + b * self.deg[i]
                + (pair - self.node_ptr[i]))

    def csr(self):
        """(indptr, indices) of the gdof x gdof structural pattern (component numbering:
        the slots are already in CSR order)."""
        if self.numbering != 'component':
            raise ValueError('csr() is the slot order of the component numbering; use to_scipy()')
        nd, N = self.nd, self.N
        indptr = np.zeros(nd * N + 1, dtype=np.int64)
        rowlen = nd * self.deg
        indptr[1:] = np.cumsum(np.tile(rowlen, nd))
        indices = np.empty(self.nnz, dtype=np.int64)
        for a in range(nd):
            for b in range(nd):
                pair = np.arange(self.nnz_node)
                indices[self.slot(a, b, pair)] = b * N + self.pair_col
        return indptr, indices

    def to_scipy(self, values):
        """SciPy CSR of slot values (all structural slots kept)."""
        import scipy.sparse as sp
        if self.numbering == 'component':
            indptr, indices = self.csr()
            return sp.csr_matrix((values, indices, indptr), shape=(self.gdof, self.gdof))
        rows, cols = self.slot_rows_cols()
        return sp.csr_matrix((values, (rows, cols)), shape=(self.gdof, self.gdof))

    def slot_rows_cols(self):
        """Row and column dof of every K slot."""
        nd = self.nd
        rows = np.empty(self.nnz, dtype=np.int64)
        cols = np.empty(self.nnz, dtype=np.int64)
        pair = np.arange(self.nnz_node)
        for a in range(nd):
            for b in range(nd):
                s = self.slot(a, b, pair)
                rows[s] = self.dof(a, self.pair_row)
                cols[s] = self.dof(b, self.pair_col)
        return rows, cols


def union_find_merge(gdof, groups):
    """Representative map rep[d] for the Bloch merge (``sweep.py:97-144``).

    Union-find with path halving; a slave's root is attached under its master's root,
    so the first-listed master of a chain stays the representative.
    Returns (rep, slaves, masters) with slaves sorted ascending.
    """
    parent = np.arange(gdof, dtype=np.int64)

    def root(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    members = []
    for g in groups:
        g = np.asarray(g, dtype=np.int64).ravel()
        if g.size == 0:
            continue
        members.extend(int(x) for x in g)
        m = int(g[0])
        for s in g[1:]:
            rs, rm = root(int(s)), root(m)
            if rs != rm:
                parent[rs] = rm
    rep = np.arange(gdof, dtype=np.int64)
    for x in set(members):
        rep[x] = root(x)
    slaves = np.flatnonzero(rep != np.arange(gdof))
    return rep, slaves, rep[slaves]


class BlochLayout:
    """Active-set reduction ``A_aa = (Tᵀ A T)[active, active]`` as gathers over K slots.

    Follows ``FrequencySolver.__init__`` (``sweep.py:283-314``) and ``solve``
    (``sweep.py:335-364``):

    * ``active`` = all dofs minus fix ∪ slaves ∪ zheng ∪ fu, ascending;
    * ``aa``: A_aa slot <- K slots whose merged (row, col) are both active;
    * ``rhs``: active row p <- K slots with merged row active[p] and merged column in
      zheng (the ``Km @ s`` column, ``sweep.py:303``);
    * ``kts``: charge weights on the *unmerged* K (``sweep.py:305``):
      ``Q = sum_p kts_a[p] x[p] + V * kts_z``, where disp[slave] = disp[master]
      and the zheng/fu values are written last (``sweep.py:356-362``).

    The merge map and the active set are set up here; the pattern and the lists over the
    K slots (millions of entries) are built by ``sawfem_bloch_layout`` (host C++).
    """

    def __init__(self, kl: KLayout, fix, zheng, fu, groups=None, rep=None):
        gdof = kl.gdof
        self.gdof = gdof
        self.zheng = np.asarray(zheng, dtype=np.int64)
        self.fu = np.asarray(fu, dtype=np.int64)
        if rep is None:
            rep, self.slaves, self.masters = union_find_merge(gdof, groups)
        else:
            rep = np.asarray(rep, dtype=np.int64)
            if rep.shape != (gdof,) or (rep[rep] != rep).any():
                raise ValueError('rep must map every dof to a representative that maps to itself')
            self.slaves = np.flatnonzero(rep != np.arange(gdof))
            self.masters = rep[self.slaves]
        self.rep = rep
        prescribed = np.unique(np.concatenate([np.asarray(fix, dtype=np.int64), self.slaves,
                                               self.zheng, self.fu]))
        self.active = np.setdiff1d(np.arange(gdof), prescribed)
        n = self.n_active = int(self.active.size)
        pos = np.full(gdof, -1, dtype=np.int64)
        pos[self.active] = np.arange(n)
        self.pos = pos
        is_z = np.zeros(gdof, dtype=bool)
        is_z[self.zheng] = True
        is_f = np.zeros(gdof, dtype=bool)
        is_f[self.fu] = True
        # A_aa pattern and the three gather lists, built natively (sawfem_bloch_layout)
        rep = np.ascontiguousarray(rep, dtype=np.int64)
        h = C.c_void_p()
        _lib.call('sawfem_bloch_layout', C.byref(h), kl.nd, kl.N, int(kl.numbering == 'interleaved'),
                  np.ascontiguousarray(kl.node_ptr, dtype=np.int64).ctypes.data,
                  np.ascontiguousarray(kl.node_idx, dtype=np.int64).ctypes.data, rep.ctypes.data, pos.ctypes.data, n,
                  is_z.view(np.uint8).ctypes.data, is_f.view(np.uint8).ctypes.data)
        a = _lib.host_arrays(h, {k: (32 + i, np.int64) for i, k in enumerate(
            ('aa_indptr', 'aa_indices', 'aa_ptr', 'aa_idx', 'rhs_ptr', 'rhs_idx', 'kts_ptr', 'kts_idx'))})
        self.aa_indptr, self.aa_indices = a['aa_indptr'], a['aa_indices']
        self.nnz_aa = int(self.aa_indices.size)
        self.aa = GatherList(a['aa_ptr'], a['aa_idx'])
        self.rhs = GatherList(a['rhs_ptr'], a['rhs_idx'])
        self.kts = GatherList(a['kts_ptr'], a['kts_idx'])
        # merged-node view of the active unknowns (supervariables for the ordering)
        self.node_of, self.comp_of = kl.node_comp(self.active)

    def expand(self, x, V=1.0):
        """Full-length displacement/potential vector from the active solution (``sweep.py:356-360``)."""
        disp = np.zeros(self.gdof, dtype=np.complex128)
        disp[self.active] = x
        disp[self.slaves] = disp[self.masters]
        disp[self.zheng] = V
        disp[self.fu] = 0.0
        return disp


def residual_groups(indptr, indices, sv_of, gmax=4):
    """Row groups for the grouped residual (``sawfem_residual_sv``): the rows of one
    supervariable (``sv_of``, e.g. the active unknowns of one merged mesh node) in groups of at
    most ``gmax`` whose column lists are identical; a row whose list differs from its group's
    first row's is split off into a group of its own.  Returns (gptr int32 [ng + 1], grows
    int32 [n]): group g holds rows grows[gptr[g]:gptr[g + 1]], every row exactly once."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices)
    sv_of = np.asarray(sv_of)
    n = indptr.size - 1
    order = np.argsort(sv_of, kind='stable').astype(np.int64)
    sv = sv_of[order]
    pos = np.arange(n)
    run_start = np.flatnonzero(np.r_[True, sv[1:] != sv[:-1]])
    run_id = np.cumsum(np.r_[True, sv[1:] != sv[:-1]]) - 1
    rank = pos - run_start[run_id]                       # position within the supervariable
    key = run_id * n + rank // gmax                      # candidate groups of <= gmax rows
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    gid = np.cumsum(np.r_[True, key[1:] != key[:-1]]) - 1
    lead = order[start[gid]]                             # first row of each row's group
    length = indptr[order + 1] - indptr[order]
    same = length == (indptr[lead + 1] - indptr[lead])
    # entries: row r's k-th column against its leader's k-th column
    cand = np.flatnonzero(same & (order != lead))
    if cand.size:
        ln = length[cand]
        rows_e = np.repeat(cand, ln)
        k = np.arange(ln.sum()) - np.repeat(np.cumsum(ln) - ln, ln)
        eq = indices[indptr[order[rows_e]] + k] == indices[indptr[lead[rows_e]] + k]
        bad = np.unique(rows_e[~eq])
        same[bad] = False
    # rows that differ from their leader become groups of their own (placed after the others)
    keep = same | (order == lead)
    rows_kept = order[keep]
    gk = gid[keep]
    counts = np.bincount(gk, minlength=gid.max() + 1 if n else 0)
    counts = counts[counts > 0]
    loners = order[~keep]
    grows = np.concatenate([rows_kept, loners]).astype(np.int32)
    sizes = np.concatenate([counts, np.ones(loners.size, dtype=np.int64)])
    gptr = np.r_[0, np.cumsum(sizes)].astype(np.int32)
    return gptr, grows
