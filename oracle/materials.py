"""Oracle: material tensors, Euler rotation, plane projections (test infrastructure only).

Restates ``saw2d/materials.py:28-71`` (Bond-matrix rotation), ``:250-265``
(ZXZ rotation matrix, degrees converted by ``*pi/180``), ``:281-308``
(xy / xz plane-strain projections), ``:74-83`` (isotropic stiffness) and
``crystal.py:346-394`` (independent constants -> tensors).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS0_CRYSTAL = 8.8541878128e-12      # crystal.py:326
EPS0_LEGACY = 8.854e-12              # materials.py:15


@dataclass
class Material:
    rho: float
    C: np.ndarray
    e: np.ndarray
    eps: np.ndarray


def rotation_matrix(euler_deg):
    """A = Rz(alpha) Rx(beta) Rz(gamma), angles in degrees (materials.py:250-265)."""
    al, be, ga = np.asarray(euler_deg, dtype=float) * np.pi / 180.0

    def rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(be), -np.sin(be)], [0.0, np.sin(be), np.cos(be)]])
    return rz(al) @ rx @ rz(ga)


def bond_matrix(a):
    """6x6 stress-transformation (Bond) matrix for Voigt order xx,yy,zz,yz,xz,xy
    (materials.py:48-66)."""
    M = np.zeros((6, 6))
    pairs = [(1, 2), (0, 2), (0, 1)]                      # Voigt 3,4,5 -> index pairs
    for i in range(3):
        for j in range(3):
            M[i, j] = a[i, j] ** 2
        for j, (k, l) in enumerate(pairs):
            M[i, 3 + j] = 2 * a[i, k] * a[i, l]
    for i, (p, q) in enumerate(pairs):
        for j in range(3):
            M[3 + i, j] = a[p, j] * a[q, j]
        for j, (k, l) in enumerate(pairs):
            M[3 + i, 3 + j] = a[p, k] * a[q, l] + a[p, l] * a[q, k]
    return M


def rotate(euler_deg, C, e, eps):
    """C' = M C Mᵀ, e' = A e Mᵀ, ε' = A ε Aᵀ (materials.py:68-70)."""
    a = rotation_matrix(euler_deg)
    M = bond_matrix(a)
    return M @ C @ M.T, a @ e @ M.T, a @ eps @ a.T


def isotropic_C(E, nu):
    """Isotropic 6x6 stiffness from E, nu (materials.py:74-83)."""
    f = E / ((1 - 2 * nu) * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = nu * f
    np.fill_diagonal(C[:3, :3], (1 - nu) * f)
    C[3, 3] = C[4, 4] = C[5, 5] = E / 2 / (1 + nu)
    return C


def crystal_tensors(symmetry, k):
    """(C, e, eps) from independent constants (crystal.py:346-394); GPa, C/m², relative ε."""
    G = 1e9
    C, e = np.zeros((6, 6)), np.zeros((3, 6))
    if symmetry == 'isotropic':
        E, nu = k['E_gpa'] * G, k['nu']
        lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        C[:3, :3] = lam
        for i in range(3):
            C[i, i] = lam + 2 * mu
            C[i + 3, i + 3] = mu
        eps = np.eye(3) * k.get('eps_r', 1.0) * EPS0_CRYSTAL
    elif symmetry == 'cubic':
        C[:3, :3] = k['C12'] * G
        for i in range(3):
            C[i, i] = k['C11'] * G
            C[i + 3, i + 3] = k['C44'] * G
        eps = np.eye(3) * k.get('eps_r', 1.0) * EPS0_CRYSTAL
    elif symmetry in ('hexagonal_6mm', 'trigonal_3m'):
        c11, c12, c13, c33, c44 = (k[x] * G for x in ('C11', 'C12', 'C13', 'C33', 'C44'))
        C[0, 0] = C[1, 1] = c11
        C[0, 1] = C[1, 0] = c12
        C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = c13
        C[2, 2] = c33
        C[3, 3] = C[4, 4] = c44
        C[5, 5] = (c11 - c12) / 2
        e[0, 4] = e[1, 3] = k['e15']
        e[2, 0] = e[2, 1] = k['e31']
        e[2, 2] = k['e33']
        if symmetry == 'trigonal_3m':
            c14 = k['C14'] * G
            C[0, 3] = C[3, 0] = c14
            C[1, 3] = C[3, 1] = -c14
            C[4, 5] = C[5, 4] = c14
            e[0, 5] = -k['e22']
            e[1, 0] = -k['e22']
            e[1, 1] = k['e22']
        eps = np.diag([k['eps11'], k['eps11'], k['eps33']]) * EPS0_CRYSTAL
    else:
        raise ValueError(symmetry)
    return C, e, eps


def project_xy(m: Material) -> Material:
    """uz = 0, Ez = 0 projection: strain [0,1,5], field [0,1] (materials.py:281-290)."""
    s, f = [0, 1, 5], [0, 1]
    return Material(m.rho, m.C[np.ix_(s, s)], m.e[np.ix_(f, s)], m.eps[np.ix_(f, f)])


def project_xz(m: Material) -> Material:
    """TC-SAW xz projection: strain [0,2,4], field [0,2] (materials.py:303-308)."""
    s, f = [0, 2, 4], [0, 2]
    return Material(m.rho, m.C[np.ix_(s, s)], m.e[np.ix_(f, s)], m.eps[np.ix_(f, f)])
