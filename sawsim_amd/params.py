"""Design parameters of the piezoelectric material and their tensor Jacobians (host side).

The 17 parameters of dY11/dp follow the reference's inputs:

* 3m independent constants C11 C12 C13 C14 C33 C44 (GPa), e15 e22 e31 e33 (C/m²),
  eps11 eps33 (relative) -> ``crystal.tensors('trigonal_3m', ...)`` (``crystal.py:371-393``);
* ZXZ Euler angles alpha beta gamma in degrees (``saw2d/materials.py:250-265``);
* Rayleigh stiffness damping beta_dK in s (``solver.py:259-265``, ``sweep.py:316-323``);
* dielectric loss tan(delta) = eta_eps, eps -> eps (1 - i eta) (``models.py:68-70``).

``lifted_tensors(p)`` gives the piezo material's kernel tensors (C9 9x9, E9 3x9, EPS 3x3,
lifted from the plane-strain projection for the 2D system).  Its Jacobian with respect to
the 15 tensor parameters is taken by the complex-step method (exact to rounding for the
real-analytic map from constants and angles to rotated tensors); the loss and damping
derivatives are closed form.
"""
from __future__ import annotations

import numpy as np

from .materials import EPS0

NAMES = ['C11', 'C12', 'C13', 'C14', 'C33', 'C44', 'e15', 'e22', 'e31', 'e33', 'eps11', 'eps33',
         'alpha', 'beta', 'gamma', 'beta_dk', 'tand']
_LIFT9 = [0, 1, 2, 3, 4, 5, 3, 4, 5]
_VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def _rot(al, be, ga):
    ca, sa, cb, sb, cg, sg = np.cos(al), np.sin(al), np.cos(be), np.sin(be), np.cos(ga), np.sin(ga)
    z = 0 * ca
    Rz1 = np.array([[ca, -sa, z], [sa, ca, z], [z, z, z + 1]])
    Rx = np.array([[z + 1, z, z], [z, cb, -sb], [z, sb, cb]])
    Rz2 = np.array([[cg, -sg, z], [sg, cg, z], [z, z, z + 1]])
    return Rz1 @ Rx @ Rz2


def _bond(a):
    """Stress transformation for Voigt order xx, yy, zz, yz, xz, xy (complex-safe)."""
    M = np.zeros((6, 6), dtype=a.dtype)
    for I, (i, j) in enumerate(_VOIGT):
        for J, (k, l) in enumerate(_VOIGT):
            M[I, J] = a[i, k] * a[j, l] + (a[i, l] * a[j, k] if k != l else 0)
    return M


def _tensors(v, plane):
    """Lossless piezo tensors from the first 15 parameters (complex-safe for the step)."""
    G = 1e9
    c11, c12, c13, c14, c33, c44 = (v[i] * G for i in range(6))
    e15, e22, e31, e33, k11, k33 = v[6:12]
    z = 0 * c11
    C = np.array([[c11, c12, c13, c14, z, z], [c12, c11, c13, -c14, z, z], [c13, c13, c33, z, z, z],
                  [c14, -c14, z, c44, z, z], [z, z, z, z, c44, c14], [z, z, z, z, c14, (c11 - c12) / 2]])
    e = np.array([[z, z, z, z, e15, -e22], [-e22, e22, z, e15, z, z], [e31, e31, e33, z, z, z]])
    eps = np.array([[k11, z, z], [z, k11, z], [z, z, k33]]) * EPS0
    a = _rot(*(np.asarray(v[12:15]) * np.pi / 180.0))
    M = _bond(a)
    C, e, eps = M @ C @ M.T, a @ e @ M.T, a @ eps @ a.T
    if plane:
        s, f = ([0, 1, 5], [0, 1]) if plane == 'xy' else ([0, 2, 4], [0, 2])
        Cp = np.zeros_like(C)
        ep = np.zeros_like(e)
        epsp = np.zeros_like(eps)
        # projection to the plane followed by the zero-filled lift used by the 2D kernel
        tgt_s, tgt_f = [0, 1, 5], [0, 1]
        Cp[np.ix_(tgt_s, tgt_s)] = C[np.ix_(s, s)]
        ep[np.ix_(tgt_f, tgt_s)] = e[np.ix_(f, s)]
        epsp[np.ix_(tgt_f, tgt_f)] = eps[np.ix_(f, f)]
        C, e, eps = Cp, ep, epsp
    return C[np.ix_(_LIFT9, _LIFT9)], e[:, _LIFT9], eps


def lifted_tensors(p, plane='xy'):
    """(C9, E9, EPS) complex128 of the piezo material for parameter vector ``p`` (17)."""
    v = np.asarray(p, dtype=float)
    C9, E9, EPS = _tensors(v, plane)
    return C9.astype(complex), E9.astype(complex), EPS * (1 - 1j * v[16])


def jacobian(p, plane='xy', h=1e-30):
    """dC9, dE9, dEPS with respect to the 17 parameters, shapes (17,9,9), (17,3,9), (17,3,3)."""
    v = np.asarray(p, dtype=float)
    dC = np.zeros((17, 9, 9), dtype=complex)
    dE = np.zeros((17, 3, 9), dtype=complex)
    dP = np.zeros((17, 3, 3), dtype=complex)
    loss = 1 - 1j * v[16]
    for k in range(15):
        w = v.astype(complex)
        w[k] += 1j * h
        C9, E9, EPS = _tensors(w, plane)
        dC[k] = C9.imag / h
        dE[k] = E9.imag / h
        dP[k] = EPS.imag / h * loss
    dP[16] = -1j * _tensors(v, plane)[2]
    return dC, dE, dP


# Hex27 slabs add the piezo density (kg/m^3) as an 18th parameter
NAMES3 = NAMES + ['rho']


def full_tensors(p):
    """Piezo (C 6x6, e 3x6, eps 3x3) complex128 of a slab cell for parameter vector ``p`` (18):
    rotated full tensors, eps with the dielectric loss (1 - i tan delta)."""
    v = np.asarray(p, dtype=float)
    C, e, eps = _full(v)
    return C.astype(complex), e.astype(complex), eps * (1 - 1j * v[16])


def _full(v):
    G = 1e9
    c11, c12, c13, c14, c33, c44 = (v[i] * G for i in range(6))
    e15, e22, e31, e33, k11, k33 = v[6:12]
    z = 0 * c11
    C = np.array([[c11, c12, c13, c14, z, z], [c12, c11, c13, -c14, z, z], [c13, c13, c33, z, z, z],
                  [c14, -c14, z, c44, z, z], [z, z, z, z, c44, c14], [z, z, z, z, c14, (c11 - c12) / 2]])
    e = np.array([[z, z, z, z, e15, -e22], [-e22, e22, z, e15, z, z], [e31, e31, e33, z, z, z]])
    eps = np.array([[k11, z, z], [z, k11, z], [z, z, k33]]) * EPS0
    a = _rot(*(np.asarray(v[12:15]) * np.pi / 180.0))
    M = _bond(a)
    return M @ C @ M.T, a @ e @ M.T, a @ eps @ a.T


def jacobian_full(p, h=1e-30):
    """dC (18,6,6), dE (18,3,6), dEPS (18,3,3), drho (18,) of the slab piezo (complex step for
    the 15 tensor parameters, closed form for tan delta and rho)."""
    v = np.asarray(p, dtype=float)
    dC = np.zeros((18, 6, 6), dtype=complex)
    dE = np.zeros((18, 3, 6), dtype=complex)
    dP = np.zeros((18, 3, 3), dtype=complex)
    drho = np.zeros(18)
    loss = 1 - 1j * v[16]
    for k in range(15):
        w = v[:15].astype(complex)
        w[k] += 1j * h
        C, e, eps = _full(w)
        dC[k] = C.imag / h
        dE[k] = e.imag / h
        dP[k] = eps.imag / h * loss
    dP[16] = -1j * _full(v)[2]
    drho[17] = 1.0
    return dC, dE, dP, drho


def from_constants(constants, euler, beta_dk=0.0, tand=0.0):
    return np.array([constants[k] for k in NAMES[:12]] + list(euler) + [beta_dk, tand], dtype=float)
