"""Oracle: Hex27 thin-slab piezoelectric assembly, periodic folds and solve
(test infrastructure only).

Restates the reference's 2.5D path (``saw2d/saw3d``):

* ``h27_gauss`` / ``h27_shapes``       -> ``quadrature_volume_h27`` elements.py:63-76 (rounded
  points and weights :22-25), ``local_basis_volume_h27`` :79-105 (factor table :29-38,
  ``_PERM`` :42)
* ``geometry``                         -> ``_geometry`` assembly.py:31-70 (|detJ|)
* ``b_rows``                           -> ``_b_matrix`` :81-101 (gamma_yz keeps only duy/dz, :92)
* ``stretch_alphas`` / ``c_stretch`` /
  ``e_stretch`` / ``eps_stretch``      -> :127-179 (x regions :27, region 3 stretches z)
* ``assemble``                         -> ``form_stiffness_mass_3d`` :230-308, ``_scatter_mass``
  :311-327 (KM_SCALE, K_phiu = KM_SCALE conj(K_uphi)^T, mass padded to 4N)
* ``boundary_dofs``                    -> ``boundary_idx_vect_3ds`` boundary.py:84-148
* ``fold_operator`` / ``Solver3D``     -> bc_solve.py:25-68, 121-175 (T_fb then T_lr,
  Q = sum K[zheng, :] disp, Y = +i w Q / V as in hex_solver.py:99)

Element matrices are built per element and Gauss point from the dense 6x81 strain matrix
(the reference builds the same products as sparse B^T D B over all integration points).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

KM_SCALE = 6.723363908816403e-11             # assembly.py:24
_X_REGIONS = {1, 5, 6, 7, 8, 9, 10, 11}      # assembly.py:27

_G = 0.774596669241483
_W5, _W8 = 0.555555555555556, 0.888888888888889
# raw shape functions: (coefficient, factor in xi1, xi2, xi3); m = t(t-1), p = t(t+1), z = (t+1)(t-1)
_RAW = [(1 / 8, 'mmm'), (1 / 8, 'pmm'), (1 / 8, 'ppm'), (1 / 8, 'mpm'), (1 / 8, 'mmp'), (1 / 8, 'pmp'),
        (1 / 8, 'ppp'), (1 / 8, 'mpp'), (-1 / 4, 'zmm'), (-1 / 4, 'pzm'), (-1 / 4, 'zpm'), (-1 / 4, 'mzm'),
        (-1 / 4, 'mmz'), (-1 / 4, 'pmz'), (-1 / 4, 'ppz'), (-1 / 4, 'mpz'), (-1 / 4, 'zmp'), (-1 / 4, 'pzp'),
        (-1 / 4, 'zpp'), (-1 / 4, 'mzp'), (1 / 2, 'zzm'), (1 / 2, 'zmz'), (1 / 2, 'pzz'), (1 / 2, 'zpz'),
        (1 / 2, 'mzz'), (1 / 2, 'zzp'), (-1.0, 'zzz')]
_ORDER = [6, 5, 4, 7, 2, 1, 0, 3, 17, 18, 14, 16, 13, 19, 12, 15, 9, 10, 8, 11, 25, 22, 23, 21, 24, 20, 26]


def _f(c, t):
    return t * (t - 1.0) if c == 'm' else (t * (t + 1.0) if c == 'p' else (t + 1.0) * (t - 1.0))


def _df(c, t):
    return 2.0 * t - 1.0 if c == 'm' else (2.0 * t + 1.0 if c == 'p' else 2.0 * t)


def h27_gauss():
    """(xi (27, 3), w (27,)) with xi1 fastest."""
    g, w = [-_G, 0.0, _G], [_W5, _W8, _W5]
    pts, wts = [], []
    for k3 in range(3):
        for k2 in range(3):
            for k1 in range(3):
                pts.append((g[k1], g[k2], g[k3]))
                wts.append(w[k1] * w[k2] * w[k3])
    return np.array(pts), np.array(wts)


def h27_shapes(xi):
    """N (q, 27) and dN (q, 27, 3) at points xi (q, 3), reference local node order."""
    xi = np.atleast_2d(np.asarray(xi, dtype=float))
    q = xi.shape[0]
    N = np.empty((q, 27))
    dN = np.empty((q, 27, 3))
    for k, (coef, fac) in enumerate(_RAW):
        a = [_f(fac[d], xi[:, d]) for d in range(3)]
        da = [_df(fac[d], xi[:, d]) for d in range(3)]
        N[:, k] = coef * a[0] * a[1] * a[2]
        dN[:, k, 0] = coef * da[0] * a[1] * a[2]
        dN[:, k, 1] = coef * a[0] * da[1] * a[2]
        dN[:, k, 2] = coef * a[0] * a[1] * da[2]
    return N[:, _ORDER], dN[:, _ORDER]


def geometry(coord, hexas):
    """Physical gradients g (ne, 27q, 27, 3) and weights |detJ| w (ne, 27q)."""
    xi, wq = h27_gauss()
    _, dN = h27_shapes(xi)
    X = coord[hexas]                                   # (ne, 27, 3)
    J = np.einsum('qai,eaj->eqij', dN, X)              # J[i][j] = d x_j / d xi_i
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                            # Jinv[j][i]: d xi_i / d x_j -> g_j = sum_i Jinv[j, i] dN_i
    g = np.einsum('eqji,qai->eqaj', Jinv, dN)
    return g, np.abs(det) * wq[None, :]


def b_rows(g):
    """Strain matrix (ne, q, 6, 81) from gradients g (ne, q, 27, 3); Voigt xx, yy, zz, yz, xz, xy;
    column 3a + c is component c of local node a; gamma_yz has no duz/dy term."""
    ne, q = g.shape[:2]
    B = np.zeros((ne, q, 6, 27, 3))
    gx, gy, gz = g[..., 0], g[..., 1], g[..., 2]
    B[:, :, 0, :, 0] = gx
    B[:, :, 1, :, 1] = gy
    B[:, :, 2, :, 2] = gz
    B[:, :, 3, :, 1] = gz                 # duy/dz (duz/dy dropped)
    B[:, :, 4, :, 0] = gz
    B[:, :, 4, :, 2] = gx
    B[:, :, 5, :, 0] = gy
    B[:, :, 5, :, 1] = gx
    return B.reshape(ne, q, 6, 81)


def stretch_alphas(region, centre, pml, omega):
    """(a1, a2, a3) per element from its centre-node coordinates (assembly.py:127-143)."""
    one = np.ones(centre.shape[0], dtype=complex)
    a1, a2, a3 = one.copy(), one.copy(), one.copy()
    if region in _X_REGIONS or region in (2, 4):
        d = pml['dmax'] * (1.0 - (centre[:, 0] - pml['xp']) ** 2 / (pml['xa'] - pml['xp']) ** 2) ** pml['n']
        a1 = 1.0 + d / (1j * omega)
    if region == 3 or region in (2, 4):
        d = pml['dmax'] * (1.0 - (centre[:, 2] - pml['zp']) ** 2 / (pml['za'] - pml['zp']) ** 2) ** pml['n']
        a3 = 1.0 + d / (1j * omega)
    return a1, a2, a3


def _ratios(a1, a2, a3):
    return a2 * a3 / a1, a1 * a3 / a2, a1 * a2 / a3


def c_stretch(a1, a2, a3):
    r1, r2, r3 = _ratios(a1, a2, a3)
    row0 = np.stack([r1, a3, a2, a3, r1, r1], -1)
    row1 = np.stack([a3, r2, a1, r2, a3, a3], -1)
    row2 = np.stack([a2, a1, r3, a1, a2, a2], -1)
    return np.stack([row0, row1, row2, row2, row2, row1], -2)


def e_stretch(a1, a2, a3):
    r1, r2, r3 = _ratios(a1, a2, a3)
    return np.stack([np.stack([r1, a3, a2, a2, a2, a3], -1), np.stack([a3, r2, a1, a1, a1, r2], -1),
                     np.stack([a2, a1, r3, r3, r3, a1], -1)], -2)


def eps_stretch(a1, a2, a3):
    r1, r2, r3 = _ratios(a1, a2, a3)
    return np.stack([np.stack([r1, a3, a2], -1), np.stack([a3, r2, a1], -1), np.stack([a2, a1, r3], -1)], -2)


def element_tensors(hexas, coord, materials, pml_list, omega):
    """Per-element C (ne,6,6), e (ne,3,6), eps (ne,3,3), rho (ne,) complex: material records
    (dicts rho, C, e, eps, indx) and PML records (region, indx, stretch parameters, own
    material) as in ``_pml_fields`` / ``_v_el`` (assembly.py:182-226)."""
    ne = hexas.shape[0]
    C = np.zeros((ne, 6, 6), dtype=complex)
    e = np.zeros((ne, 3, 6), dtype=complex)
    eps = np.zeros((ne, 3, 3), dtype=complex)
    rho = np.zeros(ne, dtype=complex)
    for pml in pml_list:
        ids = np.asarray(pml.get('indx', []), dtype=np.int64)
        if ids.size == 0:
            continue
        a1, a2, a3 = stretch_alphas(pml['region'], coord[hexas[ids, 26]], pml, omega)
        C[ids] += np.asarray(pml['C'])[None] * c_stretch(a1, a2, a3)
        e[ids] += np.asarray(pml['e'])[None] * e_stretch(a1, a2, a3)
        eps[ids] += np.asarray(pml['eps'])[None] * eps_stretch(a1, a2, a3)
        rho[ids] += pml['rho'] * a1 * a2 * a3
    for mat in materials:
        ids = np.asarray(mat['indx'], dtype=np.int64)
        C[ids] += np.asarray(mat['C'])
        e[ids] += np.asarray(mat['e'])
        eps[ids] += np.asarray(mat['eps'])
        rho[ids] += mat['rho']
    return C, e, eps, rho


def element_matrices(hexas, coord, C, e, eps, rho):
    """Kuu (ne,81,81), Kuf (ne,81,27), Kff (ne,27,27), Me (ne,27,27) before scaling."""
    g, wdet = geometry(coord, hexas)
    B = b_rows(g)                                              # (ne, q, 6, 81)
    G = np.swapaxes(g, 2, 3)                                   # (ne, q, 3, 27) potential gradients
    Kuu = np.einsum('eq,eqsa,est,eqtb->eab', wdet, B, C, B, optimize=True)
    Kuf = np.einsum('eq,eqsa,eks,eqkb->eab', wdet, B, e, G, optimize=True)
    Kff = np.einsum('eq,eqja,ejk,eqkb->eab', wdet, G, eps, G, optimize=True)
    xi, _ = h27_gauss()
    N, _ = h27_shapes(xi)
    Me = np.einsum('iq,jq,eq->eij', N.T, N.T, wdet * rho[:, None], optimize=True)
    return Kuu, Kuf, Kff, Me


def assemble(hexas, coord, materials, pml_list, omega):
    """(M, K) as (4N, 4N) complex CSR, DOFs u: 3 n + c, phi: 3 N + n (form_stiffness_mass_3d)."""
    hexas = np.asarray(hexas, dtype=np.int64)
    n_n = coord.shape[0]
    C, e, eps, rho = element_tensors(hexas, coord, materials, pml_list, omega)
    Kuu, Kuf, Kff, Me = element_matrices(hexas, coord, C, e, eps, rho)
    ne = hexas.shape[0]
    ud = (3 * hexas[:, :, None] + np.arange(3)[None, None, :]).reshape(ne, 81)
    fd = hexas
    ri = lambda a, b: np.repeat(a[:, :, None], b.shape[1], axis=2).ravel()
    ci = lambda a, b: np.repeat(b[:, None, :], a.shape[1], axis=1).ravel()
    kuu = sp.coo_matrix((Kuu.ravel(), (ri(ud, ud), ci(ud, ud))), shape=(3 * n_n, 3 * n_n)).tocsr()
    kuf = sp.coo_matrix((Kuf.ravel(), (ri(ud, fd), ci(ud, fd))), shape=(3 * n_n, n_n)).tocsr()
    kff = sp.coo_matrix((Kff.ravel(), (ri(fd, fd), ci(fd, fd))), shape=(n_n, n_n)).tocsr()
    K = sp.bmat([[kuu * KM_SCALE, kuf], [kuf.conj().T * KM_SCALE, -kff]], format='csr')
    rows, cols, vals = [], [], []
    for c in range(3):
        rows.append(ri(3 * fd + c, fd))
        cols.append(ci(fd, 3 * fd + c))
        vals.append(Me.ravel())
    M = sp.coo_matrix((np.concatenate(vals) * KM_SCALE, (np.concatenate(rows), np.concatenate(cols))),
                      shape=(4 * n_n, 4 * n_n)).tocsr()
    return M, K


# ------------------------------------------------------------------ boundary and folds
def _lr_sort(nodes, cor):
    return nodes[np.lexsort((cor[nodes, 2], cor[nodes, 1]))]


def _fb_sort(nodes, cor):
    return nodes[np.lexsort((cor[nodes, 2], cor[nodes, 0]))]


def _mech(nodes):
    nodes = np.asarray(nodes, dtype=np.int64)
    return np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])


def boundary_dofs(sets, coord, gdofu):
    """Periodic pairs and prescribed sets (boundary_idx_vect_3ds, boundary.py:84-148)."""
    i1, d1 = np.intersect1d, np.setdiff1d
    g2, L, R = sets['bottom'], sets['L'], sets['R']
    front, back, zheng, fu = sets['front'], sets['back'], sets['zheng'], sets['fu']
    fz, bz, ff, bf = i1(front, zheng), i1(back, zheng), i1(front, fu), i1(back, fu)
    front = d1(front, np.concatenate([fz, ff]))
    back = d1(back, np.concatenate([bz, bf]))
    g2F, g2b, g2L, g2R = i1(g2, front), i1(g2, back), i1(g2, L), i1(g2, R)
    g2Lb, g2Rb = i1(g2, i1(L, back)), i1(g2, i1(R, back))
    g2Le, g2Re = d1(g2L, g2Lb), d1(g2R, g2Rb)
    front, back, L, R = d1(front, g2F), d1(back, g2b), d1(L, g2L), d1(R, g2R)
    L, R = d1(L, i1(L, back)), d1(R, i1(R, back))
    cor = np.round(coord * 1e9, 1)
    L, R = _lr_sort(L, cor), _lr_sort(R, cor)
    front, back = _fb_sort(front, cor), _fb_sort(back, cor)
    g2F, g2b = _fb_sort(g2F, cor), _fb_sort(g2b, cor)
    g2Le, g2Re = _lr_sort(g2Le, cor), _lr_sort(g2Re, cor)
    fz, bz = _fb_sort(fz, cor), _fb_sort(bz, cor)
    ff, bf = _fb_sort(ff, cor), _fb_sort(bf, cor)
    m = dict(period_F=np.concatenate([_mech(front), _mech(fz), _mech(ff)]),
             period_B=np.concatenate([_mech(back), _mech(bz), _mech(bf)]),
             period_L=_mech(L), period_R=_mech(R), fix=_mech(g2))
    el = dict(period_F=np.concatenate([front, g2F]) + gdofu, period_B=np.concatenate([back, g2b]) + gdofu,
              period_L=np.concatenate([L, g2Le]) + gdofu, period_R=np.concatenate([R, g2Re]) + gdofu,
              zheng=np.asarray(zheng) + gdofu, fu=np.asarray(fu) + gdofu)
    return m, el


def fold_operator(slaves, masters, gdof):
    """0/1 matrix: slave dof takes its master's value, the others keep their own (bc_solve.py:25-33)."""
    is_slave = np.zeros(gdof, dtype=bool)
    is_slave[slaves] = True
    free = np.flatnonzero(~is_slave)
    rows = np.concatenate([free, slaves])
    cols = np.concatenate([free, masters])
    return sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(gdof, gdof)).tocsr()


class Solver3D:
    """Two-stage fold T_fb then T_lr, active set and the SuperLU solve (bc_solve.py:121-175)."""

    def __init__(self, K, M, mech, elec):
        self.K, self.M = K.tocsr(), M.tocsr()
        self.gdof = K.shape[0]
        self.mech, self.elec = mech, elec
        self.t_fb = fold_operator(np.concatenate([mech['period_B'], elec['period_B']]),
                                  np.concatenate([mech['period_F'], elec['period_F']]), self.gdof)
        self.t_lr = fold_operator(np.concatenate([mech['period_R'], elec['period_R']]),
                                  np.concatenate([mech['period_L'], elec['period_L']]), self.gdof)
        fold = lambda A: (self.t_lr.T @ (self.t_fb.T @ A @ self.t_fb) @ self.t_lr).tocsr()
        kf, mf = fold(self.K), fold(self.M)
        prescribed = np.unique(np.concatenate([mech['fix'], mech['period_B'], mech['period_R'], elec['period_R'],
                                               elec['period_B'], elec['zheng'], elec['fu']]))
        self.active = np.setdiff1d(np.arange(self.gdof), prescribed)
        self.zheng = elec['zheng']
        a, z = self.active, self.zheng
        self.k_aa, self.m_aa = kf[a][:, a].tocsc(), mf[a][:, a].tocsc()
        self.k_az, self.m_az = kf[a][:, z], mf[a][:, z]
        self.kz_rows = self.K[z, :].tocsr()

    def expand(self, x, V=1.0):
        """Full vector: active values, slaves from masters (R <- L, then B <- F), zheng = V, fu = 0."""
        m, el = self.mech, self.elec
        disp = np.zeros(self.gdof, dtype=complex)
        disp[self.active] = x
        disp[m['period_R']] = disp[m['period_L']]
        disp[m['period_B']] = disp[m['period_F']]
        disp[el['period_R']] = disp[el['period_L']]
        disp[el['period_B']] = disp[el['period_F']]
        disp[el['zheng']] = V
        disp[el['fu']] = 0.0
        return disp

    def solve_complex(self, omega, V=1.0):
        """The complex system solved directly (the reference's SuperLU fallback, bc_solve.py:160)."""
        return self.solve(omega, V)

    def solve(self, omega, V=1.0):
        """(disp, Q, Y) with Y = +i w Q / V (hex_solver.py:99)."""
        a_aa = (self.k_aa - omega ** 2 * self.m_aa).tocsc()
        a_az = self.k_az - omega ** 2 * self.m_az
        rhs = -V * np.asarray(a_az.sum(axis=1)).ravel()
        x = spla.splu(a_aa).solve(rhs)
        disp = self.expand(x, V)
        Q = (self.kz_rows @ disp).sum()
        return disp, Q, 1j * omega * Q / V
