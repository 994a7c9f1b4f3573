"""Cells matching the golden fixtures written by tests/golden/make_golden.py."""
from pathlib import Path

import numpy as np

from sawsim_amd import materials as mt
from sawsim_amd.mesh import CellMesh
from sawsim_amd.piezo_fem import Model, PMLRegion
from sawsim_amd.workloads import Cell, _al

GOLD = Path(__file__).parent / 'golden'
NAMES = ['C11', 'C12', 'C13', 'C14', 'C33', 'C44', 'e15', 'e22', 'e31', 'e33', 'eps11', 'eps33',
         'alpha', 'beta', 'gamma', 'beta_dk', 'tand']


def materials_from_params(p):
    """Projected (xy) materials [piezo, Al, Si, SiO2] for a parameter dict of the 17 names."""
    k = {key: p[key] for key in NAMES[:12]}
    pz = mt.trigonal_3m(k, 7450.0)
    C, e, eps = mt.rotate((p['alpha'], p['beta'], p['gamma']), pz.C, pz.e, pz.eps)
    pz = mt.project(mt.Material(7450.0, C, e, eps), 'xy')
    if p['tand']:
        pz = mt.Material(pz.rho, pz.C, pz.e, pz.eps * (1 - 1j * p['tand']))
    si = mt.isotropic(130e9, 0.28, 2330.0, 11.7)
    sio2 = mt.isotropic(72e9, 0.17, 2200.0, 4.2)
    return [pz] + [mt.project(m, 'xy') for m in (_al(), si, sio2)]


def tiny_cell():
    g = np.load(GOLD / 'golden_tiny_me0.npz')
    p = dict(zip(NAMES, g['params']))
    mesh = CellMesh(pos_m=g['pos_m'], quads=g['quads'], elem_tag=g['elem_tag'], lines=g['lines'],
                    info=dict(tag_pml=int(g['elem_tag'].max())))
    model = Model(mesh.pos_m, mesh.quads, mesh.elem_tag, mesh.N)
    bottom = float(mesh.pos_m[:, 1].min())
    tag_pml = 22 + 1 + 1     # one backing (22), one coating (23), PML 24
    pml = [PMLRegion(tag=tag_pml, kind='y', yp=bottom, ya=bottom + 1.085e-6)]
    cell = Cell(mesh=mesh, model=model, materials=materials_from_params(p), mat_idx=g['mat_idx'], pml=pml,
                ndisp=2, band_hz=(1.6e9, float(g['omega_top']) / (2 * np.pi)), points=3,
                beta_dk=p['beta_dk'], eta_eps=p['tand'])
    return cell, g, p


def cfg3_coarse_cell():
    """Config-3 cell of golden_cfg3_coarse.npz (unjittered sp_baseline constants, mesh 0.09 µm)."""
    from sawsim_amd.mesh import stack_cell, stack_tag_map
    from sawsim_amd.piezo_fem import material_index
    g = np.load(GOLD / 'golden_cfg3_coarse.npz')
    p = dict(zip(NAMES, g['params']))
    mesh = stack_cell(1.085, 0.5, 0.08 * 1.085, 0.6, [5 * 1.085 - 0.6], [], float(g['mesh_um']), pml_pitches=1.0)
    model = Model(mesh.pos_m, mesh.quads, mesh.elem_tag, mesh.N)
    bottom = float(mesh.pos_m[:, 1].min())
    pml = [PMLRegion(tag=mesh.info['tag_pml'], kind='y', yp=bottom, ya=bottom + mesh.info['pml_thickness_m'])]
    f = g['frequency_hz']
    cell = Cell(mesh=mesh, model=model, materials=materials_from_params(p), mat_idx=material_index(model, stack_tag_map(mesh)),
                pml=pml, ndisp=2, band_hz=(float(f[0]), float(f[-1])), points=len(f), beta_dk=p['beta_dk'],
                eta_eps=p['tand'])
    cell.params = np.array([p[k] for k in NAMES])
    return cell, g
