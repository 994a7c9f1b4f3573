"""Material tensors for the assembly kernels (host side).

Builds the per-material inputs of ``sawfem_elem_q9``: the 9x9 / 3x9 lifted stiffness and
piezoelectric tensors and the 3x3 permittivity, from SawSim-style records.

Conventions follow the reference:

* Voigt order xx, yy, zz, yz, xz, xy; stress-charge form (``crystal.py:346-394``);
* orientation by intrinsic ZXZ Euler angles in degrees, A = Rz(a) Rx(b) Rz(g), with
  degrees converted as ``angle * pi / 180`` (``saw2d/materials.py:250-265``);
* rotation C' = M C Mᵀ, e' = A e Mᵀ, eps' = A eps Aᵀ (``saw2d/materials.py:28-71``),
  implemented here through the 4th-rank tensor transformation;
* plane projections xy: strain (0, 1, 5), field (0, 1); xz: strain (0, 2, 4), field (0, 2)
  (``saw2d/materials.py:281-308``), lifted back by zero fill for the 2D kernel
  (``piezo_fem.py:389-401``).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS0 = 8.8541878128e-12          # crystal builders (crystal.py:326)
EPS0_LEGACY = 8.854e-12          # legacy records and saw2d/materials.py:15

_VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


@dataclass
class Material:
    rho: float
    C: np.ndarray
    e: np.ndarray
    eps: np.ndarray


def rotation_matrix(euler_deg):
    al, be, ga = np.asarray(euler_deg, dtype=float) * np.pi / 180.0
    ca, sa, cb, sb, cg, sg = np.cos(al), np.sin(al), np.cos(be), np.sin(be), np.cos(ga), np.sin(ga)
    Rz1 = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    Rx = np.array([[1.0, 0.0, 0.0], [0.0, cb, -sb], [0.0, sb, cb]])
    Rz2 = np.array([[cg, -sg, 0.0], [sg, cg, 0.0], [0.0, 0.0, 1.0]])
    return Rz1 @ Rx @ Rz2


def _to_tensor4(C):
    T = np.zeros((3, 3, 3, 3), dtype=np.asarray(C).dtype)
    for I, (i, j) in enumerate(_VOIGT):
        for J, (k, l) in enumerate(_VOIGT):
            for a, b in {(i, j), (j, i)}:
                for c, d in {(k, l), (l, k)}:
                    T[a, b, c, d] = C[I, J]
    return T


def _to_tensor3(e):
    T = np.zeros((3, 3, 3), dtype=np.asarray(e).dtype)
    for J, (k, l) in enumerate(_VOIGT):
        T[:, k, l] = e[:, J]
        T[:, l, k] = e[:, J]
    return T


def rotate(euler_deg, C, e, eps):
    """Rotate (C, e, eps) from crystal to device axes by 4th/3rd/2nd-rank transformation."""
    a = rotation_matrix(euler_deg)
    C4 = np.einsum('ip,jq,kr,ls,pqrs->ijkl', a, a, a, a, _to_tensor4(np.asarray(C)))
    e3 = np.einsum('ip,jq,kr,pqr->ijk', a, a, a, _to_tensor3(np.asarray(e)))
    Cv = np.array([[C4[i, j, k, l] for (k, l) in _VOIGT] for (i, j) in _VOIGT])
    ev = np.array([[e3[n, k, l] for (k, l) in _VOIGT] for n in range(3)])
    return Cv, ev, a @ np.asarray(eps) @ a.T


def isotropic(E, nu, rho, eps_r=1.0, eps0=EPS0):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] = lam + 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return Material(rho, C, np.zeros((3, 6)), np.eye(3) * eps_r * eps0)


def trigonal_3m(k, rho):
    """3m crystal from independent constants (GPa, C/m², relative ε); C66 = (C11 - C12)/2,
    +C14 at (0,3), -(1,3), (4,5); e22 at -(0,5), -(1,0), +(1,1) (``crystal.py:371-393``)."""
    G = 1e9
    c11, c12, c13, c14, c33, c44 = (k[x] * G for x in ('C11', 'C12', 'C13', 'C14', 'C33', 'C44'))
    C = np.array([[c11, c12, c13, c14, 0, 0], [c12, c11, c13, -c14, 0, 0], [c13, c13, c33, 0, 0, 0],
                  [c14, -c14, 0, c44, 0, 0], [0, 0, 0, 0, c44, c14], [0, 0, 0, 0, c14, (c11 - c12) / 2]])
    e15, e22, e31, e33 = k['e15'], k['e22'], k['e31'], k['e33']
    e = np.array([[0, 0, 0, 0, e15, -e22], [-e22, e22, 0, e15, 0, 0], [e31, e31, e33, 0, 0, 0]], dtype=float)
    eps = np.diag([k['eps11'], k['eps11'], k['eps33']]) * EPS0
    return Material(rho, C, e, eps)


def project(m: Material, plane='xy') -> Material:
    s, f = ([0, 1, 5], [0, 1]) if plane == 'xy' else ([0, 2, 4], [0, 2])
    return Material(m.rho, m.C[np.ix_(s, s)], m.e[np.ix_(f, s)], m.eps[np.ix_(f, f)])


def lift_projected(m: Material) -> Material:
    """Zero-filled full tensors from projected ones (uz = 0, Ez = 0 system)."""
    s, f = [0, 1, 5], [0, 1]
    C = np.zeros((6, 6), dtype=np.result_type(m.C, float))
    e = np.zeros((3, 6), dtype=np.result_type(m.e, float))
    eps = np.zeros((3, 3), dtype=np.result_type(m.eps, float))
    C[np.ix_(s, s)] = m.C
    e[np.ix_(f, s)] = m.e
    eps[np.ix_(f, f)] = m.eps
    return Material(m.rho, C, e, eps)


_LIFT9 = [0, 1, 2, 3, 4, 5, 3, 4, 5]


def kernel_tensors(materials):
    """(C9 (nmat,9,9), E9 (nmat,3,9), EPS (nmat,3,3) complex128; rho (nmat,)) for the K1 kernel."""
    C9 = np.stack([np.asarray(m.C, dtype=complex)[np.ix_(_LIFT9, _LIFT9)] for m in materials])
    E9 = np.stack([np.asarray(m.e, dtype=complex)[:, _LIFT9] for m in materials])
    EP = np.stack([np.asarray(m.eps, dtype=complex) for m in materials])
    rho = np.array([float(m.rho) for m in materials])
    return C9, E9, EP, rho
