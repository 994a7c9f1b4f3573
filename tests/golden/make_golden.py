"""Write golden vectors by running the reference SawSim modules on this build's meshes.

Run in the development container only (it imports the read-only reference checkout
given by --ref, default /root/reference/src); the resulting .npz files are data and are
what the tests read.  Nothing here is imported by the product or by GPU tests.

Covers what no closed-form or stored-curve check pins:

* ``golden_tiny_me0.npz``: K, M, D of a tiny plane-strain cell with a rotated, lossy 3m
  piezo, Al electrodes, Si backing, SiO2 coating and a bottom PML (reference
  ``assemble_ME0``), the value-dependent ``Km_aa`` pattern after SciPy's SpGEMM, the DOF
  maps (fix, zheng, fu, active, slaves, masters) and ``kts``.
* ``golden_cfg3_coarse.npz``: the reference CPU sweep (real-block SuperLU, as its fallback
  backend) of config 3 on a coarse mesh at 5 frequencies, and central finite differences of
  Y11 for the 17 parameters (relative step 1e-6; 1e-6 rad absolute for zero angles).
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))


def _ref_modules(ref):
    sys.path.insert(0, ref)
    from sawsim.saw2d import materials as rmat, piezo_fem as rfem, sweep as rsweep
    from sawsim import crystal as rcrys
    return rmat, rfem, rsweep, rcrys


def _materials(rmat, rcrys, params, ndisp=2):
    """Reference material pipeline: crystal builder, ZXZ rotation, xy projection, loss on piezo."""
    k = {key: params[key] for key in ('C11', 'C12', 'C13', 'C14', 'C33', 'C44', 'e15', 'e22', 'e31', 'e33',
                                       'eps11', 'eps33')}
    C, e, eps = (np.array(x, dtype=float) for x in rcrys.tensors('trigonal_3m', k))
    C, e, eps = rmat.oula_transfer([params['alpha'], params['beta'], params['gamma']], C, e, eps)
    pz = rmat.Material(7450.0, C, e, eps)
    al = rmat.Material(2700.0, rmat.computing_C_isotropic(0.33, 70e9), np.zeros((3, 6)), np.eye(3) * rmat.EPS0)
    si = rmat.Material(2330.0, rmat.computing_C_isotropic(0.28, 130e9), np.zeros((3, 6)),
                       np.eye(3) * 11.7 * rcrys.EPS0)
    sio2 = rmat.Material(2200.0, rmat.computing_C_isotropic(0.17, 72e9), np.zeros((3, 6)),
                         np.eye(3) * 4.2 * rcrys.EPS0)
    mats = [rmat.project_xy_plane_strain(m) for m in (pz, al, si, sio2)]
    if params['tand']:
        m0 = mats[0]
        mats[0] = rmat.Material(m0.rho, m0.C, m0.e, m0.eps * (1 - 1j * params['tand']))
    return mats


def _cell(mesh_um, coatings):
    from sawsim_amd.mesh import stack_cell, stack_tag_map
    m = stack_cell(1.085, 0.5, 0.08 * 1.085, 0.6, [5 * 1.085 - 0.6], coatings, mesh_um, pml_pitches=1.0)
    tm = stack_tag_map(m)
    return m, np.array([tm[t] for t in m.elem_tag])


def _engine(rfem, rsweep, rmat, mesh, idx, mats, omega_top, beta):
    model = rfem.Model(pos_m=mesh.pos_m, quads=mesh.quads, elem_tag=mesh.elem_tag, N=mesh.N)
    bottom = float(mesh.pos_m[model.nodes_with_tag(21, mesh.lines), 1].min())
    pml = [rfem.PMLRegion(tag=mesh.info['tag_pml'], kind='y', yp=bottom, ya=bottom + mesh.info['pml_thickness_m'])]
    K, M = rfem.assemble_ME0(model, mats, idx, pml, omega_top)
    bcs = rsweep.bc_periodic(model, mesh.lines, ndisp=2)
    eng = rsweep.FrequencySolver(K, M, *bcs)
    damping = [rmat.Material(m.rho, m.C, 0 * m.e, 0 * m.eps) if i == 0 else
               rmat.Material(m.rho, 0 * m.C, 0 * m.e, 0 * m.eps) for i, m in enumerate(mats)]
    D, _ = rfem.assemble_ME0(model, damping, idx, pml, omega_top)
    if beta:
        eng.set_damping(D, beta)
    return eng, K, M, D, bcs


BASE = dict(C11=232.966, C12=46.8904, C13=80.2342, C14=-11.0267, C33=275.364, C44=93.8995, e15=2.59576,
            e22=1.58923, e31=0.0821598, e33=1.88197, eps11=40.9, eps33=43.3, alpha=0.0, beta=-42.0, gamma=0.0,
            beta_dk=1e-13, tand=1e-3)
PARAMS = ['C11', 'C12', 'C13', 'C14', 'C33', 'C44', 'e15', 'e22', 'e31', 'e33', 'eps11', 'eps33',
          'alpha', 'beta', 'gamma', 'beta_dk', 'tand']


def tiny(rmat, rfem, rsweep, rcrys):
    mesh, idx = _cell(0.3, [0.3 * 1.085])
    p = dict(BASE, alpha=7.0, gamma=-11.0)
    mats = _materials(rmat, rcrys, p)
    wtop = 2 * np.pi * 2.0e9
    eng, K, M, D, bcs = _engine(rfem, rsweep, rmat, mesh, idx, mats, wtop, p['beta_dk'])
    fix, zheng, fu, groups = bcs
    Kaa = eng.Km_aa.tocoo()
    out = dict(pos_m=mesh.pos_m, quads=mesh.quads, elem_tag=mesh.elem_tag, lines=mesh.lines, mat_idx=idx,
               omega_top=wtop, K_indptr=K.indptr, K_indices=K.indices, K_data=K.data,
               M_indptr=M.indptr, M_indices=M.indices, M_data=M.data,
               D_indptr=D.indptr, D_indices=D.indices, D_data=D.data,
               fix=fix, zheng=zheng, fu=fu, groups=np.array(groups), active=eng.active,
               slaves=eng.slaves, masters=eng.masters, kts=eng.kts,
               Kaa_row=Kaa.row, Kaa_col=Kaa.col, Kaa_data=Kaa.data, Km_s=eng.Km_s, Mm_s=eng.Mm_s,
               params=np.array([p[k] for k in PARAMS]))
    np.savez_compressed(HERE / 'golden_tiny_me0.npz', **out)
    print('tiny: N', mesh.N, 'ne', mesh.ne, 'nnz K', K.nnz, 'n_active', eng.n_active)


def cfg3(rmat, rfem, rsweep, rcrys, npts=5):
    mesh, idx = _cell(0.09, [])
    band = (1.63e9, 1.915e9)
    f = np.linspace(band[0], band[1], npts)
    wtop = 2 * np.pi * band[1]

    import scipy.sparse.linalg as spla

    def y11(p):
        """(reference solve, reference system solved by SuperLU + 3 FP64 refinement steps)"""
        mats = _materials(rmat, rcrys, p)
        eng = _engine(rfem, rsweep, rmat, mesh, idx, mats, wtop, p['beta_dk'])[0]
        q = np.array([eng.solve(2 * np.pi * fz)[1] for fz in f])
        qr = []
        for fz in f:
            w = 2 * np.pi * fz
            Ks, Ka = eng.Km_s, eng.Km_aa
            if eng.beta:
                Ks = Ks + 1j * w * eng.beta * eng.Dm_s
                Ka = Ka + 1j * w * eng.beta * eng.Dm_aa
            rhs = -(Ks - w * w * eng.Mm_s)[eng.active]
            A = (Ka - w * w * eng.Mm_aa).tocsc()
            lu = spla.splu(A)
            x = lu.solve(rhs)
            for _ in range(3):
                x = x + lu.solve(rhs - A @ x)
            disp = np.zeros(eng.gdof, dtype=complex)
            disp[eng.active] = x
            disp[eng.slaves] = disp[eng.masters]
            disp[eng.zheng] = 1.0
            disp[eng.fu] = 0.0
            qr.append(eng.kts @ disp)
        return -1j * 2 * np.pi * f * q, -1j * 2 * np.pi * f * np.array(qr)
    y0, y0r = y11(BASE)
    grads = np.zeros((len(PARAMS), npts), dtype=complex)
    gradr = np.zeros((len(PARAMS), npts), dtype=complex)
    steps = np.zeros(len(PARAMS))
    for i, key in enumerate(PARAMS):
        h = 1e-6 * abs(BASE[key]) if BASE[key] != 0 else 1e-6 * (180 / np.pi)   # 1e-6 rad for zero angles
        steps[i] = h
        yp, ypr = y11(dict(BASE, **{key: BASE[key] + h}))
        ym, ymr = y11(dict(BASE, **{key: BASE[key] - h}))
        grads[i] = (yp - ym) / (2 * h)
        gradr[i] = (ypr - ymr) / (2 * h)
        print('fd', key, grads[i][2], gradr[i][2], flush=True)
    np.savez_compressed(HERE / 'golden_cfg3_coarse.npz', frequency_hz=f, y11=y0, y11_refined=y0r, dy11=grads,
                        dy11_refined=gradr, steps=steps, params=np.array([BASE[k] for k in PARAMS]),
                        names=np.array(PARAMS), mesh_um=0.09)
    print('cfg3: N', mesh.N, 'ne', mesh.ne)


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--ref', default='/root/reference/src')
    ap.add_argument('--what', default='tiny,cfg3')
    a = ap.parse_args()
    mods = _ref_modules(a.ref)
    if 'tiny' in a.what:
        tiny(*mods)
    if 'cfg3' in a.what:
        cfg3(*mods)
