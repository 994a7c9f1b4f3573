"""GPU assembly and frequency solve of the Hex27 thin slab (drop-in for SawSim ``saw2d/saw3d``).

Same entry points and argument meaning as the reference:

* ``form_stiffness_mass_3d(elem, coord, material_list, pml_list, omega) -> (mass, stiffness)``
  (``saw3d/assembly.py:230-308``): ``material_list`` entries are dicts ``rho, C (6x6), e (3x6),
  eps (3x3), indx``; ``pml_list`` entries add ``region, za, zp, xa, xp, dmax, n`` and carry their
  own material (``hex_models.py:197-202``).  K and M come back as :class:`DeviceCSR` on the
  27-node structural pattern, DOFs u: 3n + c, phi: 3N + n;
* ``find_saw_boundary_3ds`` / ``boundary_idx_vect_3ds`` (``saw3d/boundary.py:37-148``) ->
  :class:`Boundary` with ``machine`` and ``electric`` index sets;
* ``build_periodic_operators`` / ``prescribed_dofs`` / ``apply_periodic_values``
  (``saw3d/bc_solve.py:36-68``);
* ``FrequencySolver3D(K, M, boundary, t_fb, t_lr)`` with ``solve(omega, V) -> (disp, Q, |Y|)``
  and ``sweep(fre, V) -> (Q, |Y|)`` (``bc_solve.py:108-200``), Y = +i w Q / V.

The two-stage fold ``T_lrᵀ (T_fbᵀ A T_fb) T_lr`` (``bc_solve.py:10-16``) has one 1 per row of
the composite ``T_fb T_lr``, so it is the representative map ``rep = lr(fb(d))`` and the
folded, reduced system is the same index-map gather as in the 2D path.  Element matrices
come from ``sawfem_elem_h27`` and the scatter from the atomic-free ``sawfem_gather_h27``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .layout import KLayout
from .piezo_fem import DeviceCSR, _torch, stream_ptr, timed_call

KM_SCALE = 6.723363908816403e-11            # conditioning scale of the u rows (assembly.py:24)
_X_REGIONS = {1, 5, 6, 7, 8, 9, 10, 11}     # PML regions that stretch x (assembly.py:27)


@dataclass
class Boundary:
    machine: dict = field(default_factory=dict)
    electric: dict = field(default_factory=dict)


def find_saw_boundary_3ds(mesh) -> dict:
    """Boundary node sets L, R, front, back, bottom, zheng, fu of a slab mesh."""
    return {k: np.asarray(v, dtype=np.int64) for k, v in mesh.sets.items()}


def _mach(nodes):
    nodes = np.asarray(nodes, dtype=np.int64)
    return np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])


def boundary_idx_vect_3ds(sets, node_coords, gdofu) -> Boundary:
    """Periodic pairs and prescribed DOF sets of the SP slab (``boundary.py:84-148``).

    Electrode nodes leave front/back (their displacements pair separately), bottom nodes leave
    every periodic set (fixed displacements; their potentials pair front/back and left/right
    except on the back-left/right edges), L and R lose their back edges (paired through
    front/back), and each side is ordered by its rounded coordinates so pairs share an index.
    """
    i1, d1 = np.intersect1d, np.setdiff1d
    g2, L, R = sets['bottom'], sets['L'], sets['R']
    front, back, zheng, fu = sets['front'], sets['back'], sets['zheng'], sets['fu']
    fz, bz, ff, bf = i1(front, zheng), i1(back, zheng), i1(front, fu), i1(back, fu)
    front = d1(front, np.concatenate([fz, ff]))
    back = d1(back, np.concatenate([bz, bf]))
    g2F, g2b, g2L, g2R = i1(g2, front), i1(g2, back), i1(g2, L), i1(g2, R)
    g2Le = d1(g2L, i1(g2, i1(L, back)))
    g2Re = d1(g2R, i1(g2, i1(R, back)))
    front, back, L, R = d1(front, g2F), d1(back, g2b), d1(L, g2L), d1(R, g2R)
    L, R = d1(L, i1(L, back)), d1(R, i1(R, back))
    key = np.round(np.asarray(node_coords) * 1e9, 1)
    lr = lambda n: n[np.lexsort((key[n, 2], key[n, 1]))]          # by y, then z
    fb = lambda n: n[np.lexsort((key[n, 2], key[n, 0]))]          # by x, then z
    L, R, front, back = lr(L), lr(R), fb(front), fb(back)
    g2F, g2b, g2Le, g2Re = fb(g2F), fb(g2b), lr(g2Le), lr(g2Re)
    fz, bz, ff, bf = fb(fz), fb(bz), fb(ff), fb(bf)
    for a, b in ((L, R), (front, back), (g2F, g2b), (g2Le, g2Re), (fz, bz), (ff, bf)):
        if a.size != b.size:
            raise ValueError('periodic boundary node counts differ')
    bnd = Boundary()
    bnd.machine = dict(period_F=np.concatenate([_mach(front), _mach(fz), _mach(ff)]),
                       period_B=np.concatenate([_mach(back), _mach(bz), _mach(bf)]),
                       period_L=_mach(L), period_R=_mach(R), fix=_mach(g2))
    bnd.electric = dict(period_F=np.concatenate([front, g2F]) + gdofu, period_B=np.concatenate([back, g2b]) + gdofu,
                        period_L=np.concatenate([L, g2Le]) + gdofu, period_R=np.concatenate([R, g2Re]) + gdofu,
                        zheng=np.asarray(zheng, dtype=np.int64) + gdofu, fu=np.asarray(fu, dtype=np.int64) + gdofu)
    return bnd


def _fold_pairs(boundary):
    m, e = boundary.machine, boundary.electric
    fb = (np.concatenate([m['period_B'], e['period_B']]), np.concatenate([m['period_F'], e['period_F']]))
    lr = (np.concatenate([m['period_R'], e['period_R']]), np.concatenate([m['period_L'], e['period_L']]))
    return fb, lr


def build_periodic_operators(boundary, gdof):
    """(T_fb, T_lr) 0/1 fold operators (host SciPy, ``bc_solve.py:25-44``)."""
    import scipy.sparse as sp
    out = []
    for slaves, masters in _fold_pairs(boundary):
        is_slave = np.zeros(gdof, dtype=bool)
        is_slave[slaves] = True
        free = np.flatnonzero(~is_slave)
        rows = np.concatenate([free, slaves])
        cols = np.concatenate([free, masters])
        out.append(sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(gdof, gdof)))
    return tuple(out)


def fold_map(boundary, gdof):
    """Representative of every dof under T_fb then T_lr: rep = lr(fb(d))."""
    (sb, mb), (sr, mr) = _fold_pairs(boundary)
    fb = np.arange(gdof, dtype=np.int64)
    fb[sb] = mb
    lr = np.arange(gdof, dtype=np.int64)
    lr[sr] = mr
    return lr[fb]


def rep_from_operators(t_fb, t_lr):
    """Representative map of the composite fold ``T_fb T_lr`` (one 1 per row)."""
    T = (t_fb @ t_lr).tocsr()
    if not np.all(np.diff(T.indptr) == 1):
        raise ValueError('fold operators must have exactly one nonzero per row')
    return T.indices.astype(np.int64)


def prescribed_dofs(boundary):
    m, e = boundary.machine, boundary.electric
    return np.unique(np.concatenate([m['fix'], m['period_B'], m['period_R'], e['period_R'], e['period_B'],
                                     e['zheng'], e['fu']]))


def apply_periodic_values(disp, boundary, voltage):
    """Slaves from masters (R <- L, then B <- F), zheng = V, fu = 0 (``bc_solve.py:56-68``)."""
    m, e = boundary.machine, boundary.electric
    disp[m['period_R']] = disp[m['period_L']]
    disp[m['period_B']] = disp[m['period_F']]
    disp[e['period_R']] = disp[e['period_L']]
    disp[e['period_B']] = disp[e['period_F']]
    disp[e['zheng']] = voltage
    disp[e['fu']] = 0.0
    return disp


# ------------------------------------------------------------------ assembly
def _kind(region):
    return (1 if (region in _X_REGIONS or region in (2, 4)) else 0) | (2 if region in (2, 3, 4) else 0)


def material_tables(ne, material_list, pml_list):
    """Kernel inputs: per-element record and PML indices, record tensors and PML parameters."""
    recs, mat_idx, pml_idx = [], np.full(ne, -1, dtype=np.int32), np.full(ne, -1, dtype=np.int32)
    kinds, par = [], []
    for mat in material_list:
        ids = np.asarray(mat['indx'], dtype=np.int64)
        if (mat_idx[ids] >= 0).any():
            raise ValueError('an element belongs to two material records')
        mat_idx[ids] = len(recs)
        recs.append(mat)
    for pml in pml_list:
        ids = np.asarray(pml.get('indx', []), dtype=np.int64)
        if ids.size == 0:
            continue
        if (mat_idx[ids] >= 0).any():
            raise ValueError('a PML element also belongs to a material record')
        if len(kinds) == 8:
            raise ValueError('at most 8 non-empty PML records')
        mat_idx[ids] = len(recs)
        pml_idx[ids] = len(kinds)
        recs.append(pml)
        kinds.append(_kind(int(pml['region'])))
        par.append([pml['dmax'], pml['n'], pml.get('xa', 0.0), pml.get('xp', 0.0), pml.get('za', 0.0),
                    pml.get('zp', 0.0)])
    if (mat_idx < 0).any():
        raise ValueError(f'{int((mat_idx < 0).sum())} elements have no material')
    C = np.stack([np.asarray(r['C'], dtype=complex).reshape(6, 6) for r in recs])
    E = np.stack([np.asarray(r['e'], dtype=complex).reshape(3, 6) for r in recs])
    P = np.stack([np.asarray(r['eps'], dtype=complex).reshape(3, 3) for r in recs])
    rho = np.array([float(r['rho']) for r in recs])
    npml = len(kinds)
    kinds = np.array(kinds or [0], dtype=np.int32)
    par = np.ascontiguousarray(np.array(par or [[0.0] * 6], dtype=np.float64))
    return mat_idx, pml_idx, C, E, P, rho, kinds, par, npml


class _HexGeometry:
    """Device copies of the slab mesh and the 27-node gather lists."""

    def __init__(self, elem, coord):
        torch = _torch()
        dev = torch.device('cuda')
        elem = np.asarray(elem, dtype=np.int64)
        self.layout = KLayout(elem, coord.shape[0], 4, numbering='interleaved')
        kl = self.layout
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
        self.conn = t(elem.astype(np.int32))
        self.xyz = t(np.asarray(coord, dtype=np.float64))
        self.pair_row = t(kl.pair_row.astype(np.int32))
        self.node_ptr = t(kl.node_ptr.astype(np.int64))
        self.gptr = t(kl.pair_gather.ptr)
        self.gidx = t(kl.pair_gather.idx)


_HEX_CACHE = {}


def _hex_geometry(elem, coord):
    key = (id(elem), id(coord), np.shape(elem), np.shape(coord))
    g = _HEX_CACHE.get(key)
    if g is None:
        _HEX_CACHE.clear()
        g = _HEX_CACHE[key] = _HexGeometry(elem, coord)
    return g


def assemble_h27(elem, coord, material_list, pml_list, omega, damping=None):
    """K, M (and D = the K_uu part of the material records flagged in ``damping``) on the GPU."""
    torch = _torch()
    dev = torch.device('cuda')
    coord = np.asarray(coord, dtype=np.float64)
    g = _hex_geometry(elem, coord)
    kl = g.layout
    ne = kl.ne
    mat_idx, pml_idx, C, E, P, rho, kinds, par, npml = material_tables(ne, material_list, pml_list)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    dC, dE, dP, drho, dm, dpml = t(C), t(E), t(P), t(rho), t(mat_idx), t(pml_idx)
    Ke = torch.empty(ne * 16 * 729, dtype=torch.complex128, device=dev)
    Me = torch.empty(ne * 729, dtype=torch.complex128, device=dev)
    s = stream_ptr()
    timed_call('k_elem_h27', ne * (16 * 17 * 729 + 27 * (4 + 24) + 8), 'sawfem_elem_h27', s, ne, g.conn.data_ptr(),
               g.xyz.data_ptr(), dm.data_ptr(), dpml.data_ptr(), dC.data_ptr(), dE.data_ptr(), dP.data_ptr(),
               drho.data_ptr(), npml, kinds.ctypes.data, par.ctypes.data, float(omega), KM_SCALE, Ke.data_ptr(),
               Me.data_ptr())
    K = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
    M = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
    mask = np.zeros(ne, dtype=np.uint8)
    D = None
    if damping is not None:
        flags = np.asarray(damping, dtype=bool)
        for k, mat in enumerate(material_list):
            if flags[k]:
                mask[np.asarray(mat['indx'], dtype=np.int64)] = 1
        D = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
    dmask = t(mask)
    nbytes = Ke.numel() * 16 + Me.numel() * 16 + g.gidx.numel() * 8 + kl.nnz * 16 * (3 if D is not None else 2)
    timed_call('k_gather_h27', nbytes, 'sawfem_gather_h27', s, kl.nnz_node, g.pair_row.data_ptr(),
               g.node_ptr.data_ptr(), g.gptr.data_ptr(), g.gidx.data_ptr(), dmask.data_ptr(), Ke.data_ptr(),
               Me.data_ptr(), K.data_ptr(), M.data_ptr(), D.data_ptr() if D is not None else None)
    del Ke, Me
    out = [DeviceCSR(kl, K), DeviceCSR(kl, M)] + ([DeviceCSR(kl, D)] if D is not None else [])
    for mat in out:
        mat.node_xy = coord
    return tuple(out)


def form_stiffness_mass_3d(elem, coord, material_list, pml_list, omega):
    """(mass, stiffness) of the slab, the reference's return order."""
    K, M = assemble_h27(elem, coord, material_list, pml_list, omega)
    return M, K


# ------------------------------------------------------------------ solver
def _columns(coord, period):
    """Column id of every node and the (x, z) of each column: nodes at the same x (modulo the
    period) and z share their sparsity pattern in a slab one element thick, so each column is
    one supervariable of the ordering.  Coordinates are matched on a 1e-12 m grid."""
    c = np.asarray(coord, dtype=np.float64)
    kp = int(round(period * 1e12))
    kx = np.mod(np.round(c[:, 0] * 1e12).astype(np.int64), kp)
    kz = np.round(c[:, 2] * 1e12).astype(np.int64)
    uniq, col = np.unique(np.stack([kx, kz], axis=1), axis=0, return_inverse=True)
    return col.ravel(), uniq.astype(np.float64) * 1e-12


def _frequency_solver_cls():
    from .sweep import FrequencySolver
    return FrequencySolver


class FrequencySolver3D(_frequency_solver_cls()):
    """Per-frequency solve of the folded slab system on one GPU (``bc_solve.py:108-200``).

    ``t_fb``/``t_lr`` may be given as the reference builds them; otherwise the fold follows
    ``boundary`` directly.  All numeric work (batched multifrontal LU, refinement, charge) is
    the 2D engine's; the symbolic analysis orders the (x, z) columns of the slab.
    """

    def __init__(self, K: DeviceCSR, M: DeviceCSR, boundary: Boundary, t_fb=None, t_lr=None, *, period=None,
                 leaf_weight=64):
        from .layout import BlochLayout
        kl = K.layout
        if kl.numbering != 'interleaved':
            raise ValueError('FrequencySolver3D needs the Hex27 (interleaved) layout')
        rep = rep_from_operators(t_fb, t_lr) if t_fb is not None else fold_map(boundary, kl.gdof)
        self.boundary = boundary
        bl = BlochLayout(kl, boundary.machine['fix'], boundary.electric['zheng'], boundary.electric['fu'], rep=rep)
        coord = np.asarray(K.node_xy)
        if period is None:
            L, R = boundary.machine['period_L'], boundary.machine['period_R']
            period = float(abs(coord[R[0] // 3, 0] - coord[L[0] // 3, 0]))
        col, col_xz = _columns(coord, period)
        self._setup(K, M, bl, col[bl.node_of], col_xz, period, leaf_weight)

    # ---------------------------------------------------------------- gradient
    n_params = 18          # params.NAMES3: the 17 of the plane cells and the piezo density

    def enable_gradient(self, elem, coord, material_list, pml_list, omega_top, params, material=0):
        """Prepare dY11/dp for the 18 parameters (``params.NAMES3``) of material record
        ``material`` (its elements must lie outside the PML).  The other arguments are the
        assembly inputs; ``params`` is the vector the record's tensors were built from."""
        from . import params as prm
        torch = self.torch
        bl = self.bloch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=self.dev)
        elem = np.asarray(elem, dtype=np.int64)
        _, pml_idx, _, _, _, _, kinds, par, npml = material_tables(elem.shape[0], material_list, pml_list)
        elems = np.asarray(material_list[material]['indx'], dtype=np.int64)
        if (pml_idx[elems] >= 0).any():
            raise ValueError('the parameterised material must not be a PML record')
        self._grad = dict(elems=t(elems.astype(np.int32)), ne0=int(elems.size), conn=t(elem.astype(np.int32)),
                          xyz=t(np.asarray(coord, dtype=np.float64)), pml_idx=t(pml_idx), npml=npml, kinds=kinds,
                          par=par, omega_top=float(omega_top), N=int(np.asarray(coord).shape[0]),
                          pos_rep=t(bl.pos[bl.rep].astype(np.int32)),
                          zmask=t(np.isin(np.arange(self.gdof), bl.zheng).astype(np.uint8)),
                          params=np.asarray(params, dtype=float))
        self._grad['tensors'] = prm.full_tensors(params)
        self._grad['jac'] = prm.jacobian_full(params)
        self._gradient_maps(KM_SCALE)

    def _contract(self, omegas, V, x, lam):
        """dY11/dp from the forward solution x and the adjoint lambda: element sensitivities on
        the GPU (``sawfem_adjoint_h27``), parameter contraction on the host."""
        torch = self.torch
        g = self._grad
        s = stream_ptr()
        B, n = x.shape
        Bp = -(-B // 64) * 64
        xt = torch.empty((n, Bp), dtype=torch.complex128, device=self.dev)
        lt = torch.empty((n, Bp), dtype=torch.complex128, device=self.dev)
        _lib.call('sawfem_transpose', s, B, n, n, Bp, x.data_ptr(), xt.data_ptr())
        _lib.call('sawfem_transpose', s, B, n, n, Bp, lam.data_ptr(), lt.data_ptr())
        work = torch.empty(max(1, B * g['ne0'] * 64), dtype=torch.complex128, device=self.dev)
        W = torch.empty(B * 64, dtype=torch.complex128, device=self.dev)
        _lib.call('sawfem_adjoint_h27', s, g['ne0'], g['elems'].data_ptr(), g['conn'].data_ptr(), g['xyz'].data_ptr(),
                  g['pml_idx'].data_ptr(), g['npml'], g['kinds'].ctypes.data, g['par'].ctypes.data, g['omega_top'],
                  g['N'], g['pos_rep'].data_ptr(), g['zmask'].data_ptr(), float(V), KM_SCALE, B, Bp, xt.data_ptr(),
                  lt.data_ptr(), work.data_ptr(), W.data_ptr())
        Wh = W.cpu().numpy().reshape(B, 64)
        return dict(dY=self._param_gradient(Wh, omegas, V))

    def _param_gradient(self, W, omegas, V):
        """dY11/dp (B, 18) from the element sensitivity sums W (B, 64) = (Hc, He, Hp, Hm):
        dQ = (1 + i w beta) KM_SCALE Hc:dC + He:dE - Hp:dEps - w^2 KM_SCALE Hm drho, the damping
        parameter through i w KM_SCALE Hc:C, and dY = +i w dQ / V."""
        g = self._grad
        B = W.shape[0]
        Hc = W[:, :36].reshape(B, 6, 6)
        He = W[:, 36:54].reshape(B, 3, 6)
        Hp = W[:, 54:63].reshape(B, 3, 3)
        Hm = W[:, 63]
        dC, dE, dP, drho = g['jac']
        C = g['tensors'][0]
        w = np.asarray(omegas, dtype=float)[:, None]
        fac = 1 + 1j * w * self.beta
        dQ = (fac * KM_SCALE * np.einsum('bst,kst->bk', Hc, dC) + np.einsum('bks,pks->bp', He, dE)
              - np.einsum('bjk,pjk->bp', Hp, dP) - w ** 2 * KM_SCALE * Hm[:, None] * drho[None, :])
        dQ[:, 15] = (1j * w[:, 0]) * KM_SCALE * np.einsum('bst,st->b', Hc, C)
        return 1j * w / V * dQ
