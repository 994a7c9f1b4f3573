"""Shared test helpers: oracle-side objects for a workload cell."""
import numpy as np
import scipy.sparse.linalg as spla

from oracle import fem as ofem, sweep as osweep, materials as omat


def oracle_materials(cell):
    return [omat.Material(m.rho, np.asarray(m.C), np.asarray(m.e), np.asarray(m.eps)) for m in cell.materials]


def oracle_pml(cell):
    return [ofem.PMLRegion(tag=r.tag, kind=r.kind, dmax=r.dmax, n=r.n, xa=r.xa, xp=r.xp, ya=r.ya, yp=r.yp)
            for r in cell.pml]


def oracle_assemble(cell, omega_top, materials=None):
    mats = materials if materials is not None else oracle_materials(cell)
    m = cell.model
    fn = ofem.assemble if cell.ndisp == 3 else ofem.assemble_me0
    return fn(m.pos_m, m.quads, m.elem_tag, mats, cell.mat_idx, oracle_pml(cell), omega_top)


def oracle_damping_materials(cell):
    out = []
    for i, m in enumerate(oracle_materials(cell)):
        C = m.C if i == 0 else 0 * m.C
        out.append(omat.Material(m.rho, C, 0 * m.e, 0 * m.eps))
    return out


def oracle_engine(cell, omega_top):
    K, M = oracle_assemble(cell, omega_top)
    bcs = osweep.bc_periodic(cell.model.pos_m, cell.mesh.lines, cell.model.N, ndisp=cell.ndisp)
    eng = osweep.FrequencySolver(K, M, *bcs)
    if cell.beta_dk:
        D, _ = oracle_assemble(cell, omega_top, oracle_damping_materials(cell))
        eng.set_damping(D, cell.beta_dk)
    return eng


def refined_charge(eng, omega, V=1.0, steps=4):
    """Oracle Q: complex SuperLU solve plus FP64 iterative refinement with the same factors."""
    A, rhs = eng.system(omega, V)
    lu = spla.splu(A.tocsc())
    x = lu.solve(rhs)
    for _ in range(steps):
        x = x + lu.solve(rhs - A @ x)
    return eng.kts @ eng.expand(x, V), x
