"""Oracle: vectorised Q9 piezoelectric assembly (test infrastructure only).

Restates ``saw2d/piezo_fem.py``:

* ``gauss_3x3`` / ``q9_shapes``            -> ``_gauss_3x3`` :48-53, ``_q9_shapes`` :56-103
* ``lift_c9`` / ``lift_e9``                -> ``_C9`` :109-116, ``_e9`` :119-124
* ``pml_t9`` / ``pml_teps``                -> ``_T9`` :127-142, ``_Teps`` :145-155
* ``element_alpha``                        -> :222-242 (centre node, frozen at omega_top)
* ``element_geometry``                     -> :248-280 (signed detJ)
* ``assemble``                             -> :283-367 (4N lifted system, COO -> CSR)
* ``assemble_me0``                         -> :373-405 (lift, assemble, keep [ux|uy|phi])
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .materials import Material

SCALE = 6.173191780858575e-11          # piezo_fem.py:26

# displacement component -> [(row of the 9-component strain, derivative direction)]
STRAIN_SEL = {0: [(0, 0), (5, 1)], 1: [(1, 1), (8, 0)], 2: [(6, 1), (7, 0)]}   # piezo_fem.py:31-35


@dataclass
class PMLRegion:
    """piezo_fem.py:192-208"""
    tag: int
    kind: str
    dmax: float = 1e11
    n: int = 3
    xa: float = 0.0
    xp: float = 0.0
    ya: float = 0.0
    yp: float = 0.0


def gauss_3x3():
    g = np.array([-0.774596669241483, 0.0, 0.774596669241483])
    w = np.array([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])
    gx, gy = np.meshgrid(g, g)                 # x fastest, as (gx, gy) for gy in g for gx in g
    wx, wy = np.meshgrid(w, w)
    return np.column_stack([gx.ravel(), gy.ravel()]), (wx * wy).ravel()


def q9_shapes(xi, eta):
    """Biquadratic Lagrange shapes, QUADS9 order, and d/d(xi, eta).

    Built from the 1D quadratic Lagrange polynomials on nodes (-1, 0, 1):
    node a has 1D indices (i_a, j_a) with corners (-1,-1),(1,-1),(1,1),(-1,1),
    mid-edges (0,-1),(1,0),(0,1),(-1,0), centre (0,0).
    """
    xi = np.asarray(xi, float)
    eta = np.asarray(eta, float)

    def l1(t):   # values of L_{-1}, L_0, L_{+1}
        return np.stack([0.5 * t * (t - 1), 1 - t * t, 0.5 * t * (t + 1)], -1)

    def d1(t):
        return np.stack([t - 0.5, -2 * t, t + 0.5], -1)
    ia = np.array([0, 2, 2, 0, 1, 2, 1, 0, 1])
    ja = np.array([0, 0, 2, 2, 0, 1, 2, 1, 1])
    Lx, Ly, Dx, Dy = l1(xi), l1(eta), d1(xi), d1(eta)
    N = Lx[..., ia] * Ly[..., ja]
    dN = np.stack([Dx[..., ia] * Ly[..., ja], Lx[..., ia] * Dy[..., ja]], -1)
    return N, dN


def lift_c9(C):
    idx = [0, 1, 2, 3, 4, 5, 3, 4, 5]
    return np.asarray(C)[np.ix_(idx, idx)]


def lift_e9(e):
    return np.asarray(e)[:, [0, 1, 2, 3, 4, 5, 3, 4, 5]]


def pml_t9(a1, a2, a3):
    """9x9 per-element PML multipliers (piezo_fem.py:127-142)."""
    code = dict(A=a2 * a3 / a1, b=a3, c=a2, D=a1 * a3 / a2, e=a1, F=a1 * a2 / a3)
    pattern = ["AbcccbbAA", "bDeeeDDbb", "ceFFFeecc"]
    rowsel = [0, 1, 2, 2, 2, 1, 1, 0, 0]
    T = np.empty((9, 9, np.asarray(a1).shape[0]), dtype=complex)
    for r in range(9):
        for c, ch in enumerate(pattern[rowsel[r]]):
            T[r, c] = code[ch]
    return T


def pml_teps(a1, a2, a3):
    """3x3 per-element dielectric PML multipliers (piezo_fem.py:145-155)."""
    T = np.empty((3, 3, np.asarray(a1).shape[0]), dtype=complex)
    T[0] = [a2 * a3 / a1, a3, a2]
    T[1] = [a3, a1 * a3 / a2, a1]
    T[2] = [a2, a1, a1 * a2 / a3]
    return T


def element_alpha(pos_m, quads, elem_tag, pml_regions, omega):
    """Per-element stretch factors from the centre node (piezo_fem.py:222-242)."""
    ne = quads.shape[0]
    a = [np.ones(ne, dtype=complex) for _ in range(3)]
    xc, yc = pos_m[quads[:, 8], 0], pos_m[quads[:, 8], 1]
    for reg in pml_regions:
        sel = elem_tag == reg.tag
        if 'x' in reg.kind:
            r = np.clip((xc - reg.xp) ** 2 / (reg.xa - reg.xp) ** 2, 0.0, 1.0)
            a[0][sel] = 1.0 + (reg.dmax * (1.0 - r) ** reg.n)[sel] / (1j * omega)
        if 'y' in reg.kind:
            r = np.clip((yc - reg.yp) ** 2 / (reg.ya - reg.yp) ** 2, 0.0, 1.0)
            a[1][sel] = 1.0 + (reg.dmax * (1.0 - r) ** reg.n)[sel] / (1j * omega)
    return tuple(a)


def element_geometry(pos_m, quads):
    """Gram matrices G[(d1,d2)] (ne,9,9) and GM (ne,9,9) with signed detJ (:248-280)."""
    pts, wts = gauss_3x3()
    Nq, dNr = q9_shapes(pts[:, 0], pts[:, 1])
    xy = pos_m[quads]                                        # (ne,9,2)
    J = np.einsum('eai,qaj->eqij', xy, dNr)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    inv = np.stack([np.stack([J[..., 1, 1], -J[..., 0, 1]], -1),
                    np.stack([-J[..., 1, 0], J[..., 0, 0]], -1)], -2) / det[..., None, None]
    dNx = np.einsum('qaj,eqjk->eqak', dNr, inv)
    dV = wts[None, :] * det
    G = {(d1, d2): np.einsum('eq,eqa,eqb->eab', dV, dNx[..., d1], dNx[..., d2])
         for d1 in (0, 1) for d2 in (0, 1)}
    GM = np.einsum('eq,qa,qb->eab', dV, Nq, Nq)
    return G, GM


def element_blocks(pos_m, quads, elem_tag, materials, mat_idx, pml_regions, omega_top):
    """Per-element coupled blocks before scattering (assemble :297-350).

    Returns dict with Kuu[(a,b)], Kuphi[a], Kphiphi, Mblk; each (ne,9,9) complex,
    already including SCALE / sign factors except the K_phiu transpose.
    """
    G, GM = element_geometry(pos_m, quads)
    a1, a2, a3 = element_alpha(pos_m, quads, elem_tag, pml_regions, omega_top)
    T9 = pml_t9(a1, a2, a3)
    Teps = pml_teps(a1, a2, a3)
    mat_idx = np.asarray(mat_idx, dtype=int)
    C9 = np.stack([lift_c9(m.C) for m in materials])[mat_idx]          # (ne,9,9)
    E9 = np.stack([lift_e9(m.e) for m in materials])[mat_idx]          # (ne,3,9)
    EP = np.stack([np.asarray(m.eps) for m in materials])[mat_idx]     # (ne,3,3)
    rho = np.array([m.rho for m in materials])[mat_idx]
    Ceff = np.moveaxis(T9, -1, 0) * C9
    Eeff = np.moveaxis(T9[:3], -1, 0) * E9
    Peff = np.moveaxis(Teps, -1, 0) * EP
    out = dict(Kuu={}, Kuphi={})
    for a in range(3):
        for b in range(3):
            blk = np.zeros(GM.shape, dtype=complex)
            for rv, dv in STRAIN_SEL[a]:
                for ru, du in STRAIN_SEL[b]:
                    blk += Ceff[:, rv, ru, None, None] * G[(dv, du)]
            out['Kuu'][(a, b)] = blk * SCALE
    for a in range(3):
        blk = np.zeros(GM.shape, dtype=complex)
        for rv, dv in STRAIN_SEL[a]:
            for k in (0, 1):
                blk += Eeff[:, k, rv, None, None] * G[(dv, k)]
        out['Kuphi'][a] = blk
    blk = np.zeros(GM.shape, dtype=complex)
    for j in (0, 1):
        for k in (0, 1):
            blk += Peff[:, j, k, None, None] * G[(j, k)]
    out['Kphiphi'] = -blk
    out['Mblk'] = (rho * a1 * a2 * a3)[:, None, None] * GM * SCALE
    return out


def assemble(pos_m, quads, elem_tag, materials, mat_idx, pml_regions, omega_top):
    """Global K, M (4N x 4N complex CSR) with layout [ux|uy|uz|phi] (piezo_fem.py:283-367)."""
    N = pos_m.shape[0]
    ne = quads.shape[0]
    blk = element_blocks(pos_m, quads, elem_tag, materials, mat_idx, pml_regions, omega_top)
    rn = np.broadcast_to(quads[:, :, None], (ne, 9, 9))
    cn = np.broadcast_to(quads[:, None, :], (ne, 9, 9))
    R, Cc, V = [], [], []

    def put(b, ro, co):
        R.append((rn + ro).ravel())
        Cc.append((cn + co).ravel())
        V.append(b.ravel())
    for a in range(3):
        for b in range(3):
            put(blk['Kuu'][(a, b)], a * N, b * N)
    for a in range(3):
        put(blk['Kuphi'][a], a * N, 3 * N)
        put(SCALE * blk['Kuphi'][a].transpose(0, 2, 1), 3 * N, a * N)
    put(blk['Kphiphi'], 3 * N, 3 * N)
    K = sp.coo_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(Cc))),
                      shape=(4 * N, 4 * N)).tocsr()
    R, Cc, V = [], [], []
    for a in range(3):
        put(blk['Mblk'], a * N, a * N)
    M = sp.coo_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(Cc))),
                      shape=(4 * N, 4 * N)).tocsr()
    return K, M


def lift_me0(materials):
    """Zero-fill projected 3x3 / 2x3 / 2x2 tensors into 6x6 / 3x6 / 3x3 (:389-401)."""
    s, f = [0, 1, 5], [0, 1]
    out = []
    for m in materials:
        if np.shape(m.C) != (3, 3) or np.shape(m.e) != (2, 3) or np.shape(m.eps) != (2, 2):
            raise ValueError('ME0 requires projected xy tensors: C3x3, e2x3, eps2x2')
        C = np.zeros((6, 6), dtype=np.asarray(m.C).dtype)
        e = np.zeros((3, 6), dtype=np.asarray(m.e).dtype)
        eps = np.zeros((3, 3), dtype=np.asarray(m.eps).dtype)
        C[np.ix_(s, s)] = m.C
        e[np.ix_(f, s)] = m.e
        eps[np.ix_(f, f)] = m.eps
        out.append(Material(m.rho, C, e, eps))
    return out


def assemble_me0(pos_m, quads, elem_tag, materials, mat_idx, pml_regions, omega_top):
    """Plane-strain [ux|uy|phi] K, M (3N x 3N) via the lifted 4N system (:373-405)."""
    K, M = assemble(pos_m, quads, elem_tag, lift_me0(materials), mat_idx, pml_regions, omega_top)
    n = pos_m.shape[0]
    keep = np.r_[0:2 * n, 3 * n:4 * n]
    return K[keep][:, keep].tocsr(), M[keep][:, keep].tocsr()
