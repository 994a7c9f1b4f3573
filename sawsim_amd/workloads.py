"""Synthetic periodic SAW cells used by the benchmark and the parity tests.

Each builder returns a :class:`Cell`: the mesh (reference ``Model`` + line tags),
materials (already rotated and, for plane strain, projected), per-element material
index, PML regions and the sweep band.  The layer/PML/material assignment follows
``models.py:53-72`` and ``sp_meshes/__init__.py:27-41`` of the reference.

Workloads:

* ``config1``: sp_stack-style structured Q9 cell, pitch 1.085 µm, Al electrode
  0.08·pitch thick, metallization 0.5, 0.6 µm trigonal-3m piezo film (sp_baseline
  constants with seeded ±5 % jitter, Euler (0, -42°, 0)) over isotropic Si
  (E = 130 GPa, ν = 0.28, ρ = 2330), depth 5·pitch plus a 1·pitch bottom PML in Si,
  plane strain ``[ux|uy|phi]`` (mesh 0.045 µm: 4312 elements, 17521 nodes); 101 points
  over fR·[0.92, 1.08].
* ``config2``: the same stack with a 0.3·pitch SiO₂ overlay embedding the electrodes,
  Euler (0, -38°, 0), mesh refined 2x in each direction (16848 elements); 401 points over
  fR·[0.92, 1.08].
* ``config3``: config 1 with Rayleigh stiffness damping β = 1e-13 s and dielectric
  loss tanδ = 1e-3 on the piezo.
* ``double_layer``: the reference ``sp_double_layer`` template (ME1) whose |Y| curve
  the reference stores in ``tests/data/sp_double_layer.npz``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import materials as mt
from .mesh import CellMesh, double_layer_cell, quad_layer_cell, stack_cell, stack_tag_map
from .params import from_constants
from .piezo_fem import Model, PMLRegion, material_index

SP_BASELINE_3M = dict(C11=232.966, C12=46.8904, C13=80.2342, C14=-11.0267, C33=275.364, C44=93.8995,
                      e15=2.59576, e22=1.58923, e31=0.0821598, e33=1.88197, eps11=40.9, eps33=43.3)
PIEZO_KEYS = ['C11', 'C12', 'C13', 'C14', 'C33', 'C44', 'e15', 'e22', 'e31', 'e33', 'eps11', 'eps33']


@dataclass
class Cell:
    mesh: CellMesh
    model: Model
    materials: list
    mat_idx: np.ndarray
    pml: list
    ndisp: int
    band_hz: tuple
    points: int
    beta_dk: float = 0.0
    eta_eps: float = 0.0
    info: dict = field(default_factory=dict)
    params: np.ndarray = None          # 17 design parameters of the piezo material (params.NAMES)

    @property
    def nd(self):
        return self.ndisp + 1

    def frequencies(self, points=None):
        return np.linspace(self.band_hz[0], self.band_hz[1], points or self.points)


def jittered_constants(seed=0, rel=0.05):
    rng = np.random.default_rng(seed)
    return {k: v * (1.0 + rng.uniform(-rel, rel)) for k, v in SP_BASELINE_3M.items()}


def _al():
    # legacy Al record: E = 70 GPa, nu = 0.33, rho = 2700, eps = eps0 (8.854e-12) I
    return mt.isotropic(70e9, 0.33, 2700.0, 1.0, eps0=mt.EPS0_LEGACY)


def piezo_material(constants, euler_deg, eta_eps=0.0, plane='xy', rho=7450.0):
    m = mt.trigonal_3m(constants, rho)
    C, e, eps = mt.rotate(euler_deg, m.C, m.e, m.eps)
    m = mt.Material(rho, C, e, eps)
    if plane:
        m = mt.project(m, plane)
    if eta_eps:
        m = mt.Material(m.rho, m.C, m.e, m.eps * (1 - 1j * eta_eps))
    return m


def _stack(mesh, mats, plane, ndisp, band, points, beta_dk=0.0, eta_eps=0.0, info=None):
    model = Model(mesh.pos_m, mesh.quads, mesh.elem_tag, mesh.N)
    tag_map = stack_tag_map(mesh) if 'tag_backing' in mesh.info else mesh.info['tag_map']
    bottom = float(mesh.pos_m[mesh.nodes_with_tag(21), 1].min())
    pml = [PMLRegion(tag=mesh.info['tag_pml'], kind='y', yp=bottom, ya=bottom + mesh.info['pml_thickness_m'])]
    if plane:
        mats = [mats[0]] + [mt.project(m, plane) for m in mats[1:]]
    return Cell(mesh=mesh, model=model, materials=mats, mat_idx=material_index(model, tag_map), pml=pml,
                ndisp=ndisp, band_hz=band, points=points, beta_dk=beta_dk, eta_eps=eta_eps, info=info or {})


# Sweep bands: fR * [0.92, 1.08] around the main resonance fR of a 201-point pilot sweep over
# 1.0-3.0 GHz (fr of `metrics.resonances`), measured once on the default cells: config 1
# fR = 1.4286 GHz, config 2 fR = 2.5124 GHz (`pilot_resonance` recomputes them).
BAND1 = (1.3143e9, 1.5429e9)
BAND2 = (2.3114e9, 2.7134e9)


def config1(seed=0, mesh_um=0.045, points=101, band=BAND1, beta_dk=0.0, eta_eps=0.0, euler=(0, -42, 0)):
    pitch = 1.085
    mesh = stack_cell(pitch, 0.5, 0.08 * pitch, 0.6, [5 * pitch - 0.6], [], mesh_um, pml_pitches=1.0)
    k = jittered_constants(seed)
    mats = [piezo_material(k, euler, eta_eps), _al(), mt.isotropic(130e9, 0.28, 2330.0, 11.7)]
    c = _stack(mesh, mats, 'xy', 2, band, points, beta_dk, eta_eps,
               info=dict(name='config1', constants=k, euler=euler, seed=seed))
    c.params = from_constants(k, euler, beta_dk, eta_eps)
    return c


def config2(seed=0, mesh_um=0.0225, points=401, band=BAND2):
    pitch = 1.085
    mesh = stack_cell(pitch, 0.5, 0.08 * pitch, 0.6, [5 * pitch - 0.6], [0.3 * pitch], mesh_um, pml_pitches=1.0)
    k = jittered_constants(seed)
    euler = (0, -38, 0)
    mats = [piezo_material(k, euler), _al(), mt.isotropic(130e9, 0.28, 2330.0, 11.7),
            mt.isotropic(72e9, 0.17, 2200.0, 4.2)]
    c = _stack(mesh, mats, 'xy', 2, band, points, info=dict(name='config2', constants=k, euler=euler, seed=seed))
    c.params = from_constants(k, euler)
    return c


def config3(seed=0, mesh_um=0.045, points=201, band=BAND1):
    c = config1(seed, mesh_um, points, band, beta_dk=1e-13, eta_eps=1e-3)
    c.info['name'] = 'config3'
    return c


def legacy_record_materials(layers):
    """sp_baseline (rotated 0, -42, 0), Al and builtin layer records used by the stored curves."""
    C = np.array([[232966000000.0, 46890400000.0, 80234200000.0, -11026700000.0, 0, 0],
                  [46890400000.0, 232966000000.0, 80234200000.0, 11026700000.0, 0, 0],
                  [80234200000.0, 80234200000.0, 275364000000.0, 0, 0, 0],
                  [-11026700000.0, 11026700000.0, 0, 93899500000.0, 0, 0],
                  [0, 0, 0, 0, 93899500000.0, -11026700000.0],
                  [0, 0, 0, 0, -11026700000.0, 93038000000.0]])
    e = np.array([[0, 0, 0, 0, 2.59576, -1.58923], [-1.58923, 1.58923, 0, 2.59576, 0, 0],
                  [0.0821598, 0.0821598, 1.88197, 0, 0, 0]])
    eps = np.diag([3.6212859999999994e-10, 3.6212859999999994e-10, 3.833781999999999e-10])
    Cr, er, epsr = mt.rotate((0, -42, 0), C, e, eps)
    lib = dict(si_isotropic=mt.isotropic(170e9, 0.28, 2329.0, 11.7, eps0=mt.EPS0_LEGACY),
               sio2=mt.isotropic(70e9, 0.17, 2200.0, 4.2, eps0=mt.EPS0_LEGACY),
               polysi=mt.isotropic(169e9, 0.22, 2320.0, 4.5, eps0=mt.EPS0_LEGACY))
    return [mt.Material(7450.0, Cr, er, epsr), _al()] + [lib[x] for x in layers]


def double_layer(points=5, band=(1.5e9, 2.7e9)):
    mesh = double_layer_cell()
    mats = legacy_record_materials(['si_isotropic'])
    return _stack(mesh, mats, None, 3, band, points, info=dict(name='sp_double_layer'))


def quad_layer(points=5, band=(1.5e9, 2.7e9)):
    mesh = quad_layer_cell()
    mats = legacy_record_materials(['sio2', 'polysi', 'si_isotropic'])
    return _stack(mesh, mats, None, 3, band, points, info=dict(name='sp_quad_layer'))


# ------------------------------------------------------------------ design set (config 5)
def _layer_material(rng):
    """An isotropic or cubic layer record with ±5 % jitter on each independent constant."""
    j = lambda v: v * (1.0 + rng.uniform(-0.05, 0.05))
    kind = int(rng.integers(0, 4))
    if kind == 0:
        return mt.isotropic(j(130e9), j(0.28), j(2330.0), j(11.7))           # Si
    if kind == 1:
        return mt.isotropic(j(72e9), j(0.17), j(2200.0), j(4.2))             # SiO2
    if kind == 2:
        return mt.isotropic(j(169e9), j(0.22), j(2320.0), j(4.5))            # poly-Si
    C = np.zeros((6, 6))                                                      # cubic Si
    c11, c12, c44 = j(166e9), j(64e9), j(80e9)
    C[:3, :3] = c12
    C[np.arange(3), np.arange(3)] = c11
    C[np.arange(3, 6), np.arange(3, 6)] = c44
    return mt.Material(j(2330.0), C, np.zeros((3, 6)), np.eye(3) * j(11.68) * mt.EPS0)


def design_target(seed, ne_range=(2000, 20000)):
    """Element-count target of design cell ``seed`` (the last draw of :func:`design_cell`)."""
    return float(design_cell(seed, ne_range=ne_range, draws_only=True))


def design_cell(seed, points=128, ne_range=(2000, 20000), ne_target=None, draws_only=False):
    """One cell of the design set: pitch U(0.8, 2.0) µm, 0-3 backings and 0-2 coatings with
    thicknesses U(0.1, 0.8) pitch (the deepest backing fills the stack to 5 pitch; without
    backings the piezo is the substrate), isotropic/cubic layer records and a 3m piezo, all
    jittered ±5 %, Euler (0, U(-50, -35)°, 0); Q9 plane strain with a mesh size that puts the
    element count near a target drawn uniformly from ``ne_range``; Al electrodes 0.08 pitch thick;
    1 pitch bottom PML; ``points`` frequencies over [0.7, 1.3] x (3100 m/s) / (2 pitch).
    ``ne_target`` replaces the drawn element count (the same cell on a coarser or finer mesh);
    ``draws_only`` returns the drawn target without meshing."""
    rng = np.random.default_rng(seed)
    pitch = float(rng.uniform(0.8, 2.0))
    nb, nc = int(rng.integers(0, 4)), int(rng.integers(0, 3))
    thick = lambda: float(rng.uniform(0.1, 0.8)) * pitch
    if nb:
        piezo = thick()
        backs = [thick() for _ in range(nb - 1)]
        backs.append(max(0.2 * pitch, 5 * pitch - piezo - sum(backs)))
    else:
        piezo, backs = 5 * pitch, []
    coats = [thick() for _ in range(nc)]
    k = {key: v * (1.0 + rng.uniform(-0.05, 0.05)) for key, v in SP_BASELINE_3M.items()}
    euler = (0.0, float(rng.uniform(-50.0, -35.0)), 0.0)
    mats = ([piezo_material(k, euler), _al()] + [_layer_material(rng) for _ in range(nb)]
            + [_layer_material(rng) for _ in range(nc)])
    target = float(rng.uniform(*ne_range))
    if draws_only:
        return target
    if ne_target is not None:
        target = float(ne_target)
    build = lambda h: stack_cell(pitch, 0.5, 0.08 * pitch, piezo, backs, coats, h, pml_pitches=1.0)
    lo, hi = 0.005 * pitch, 0.5 * pitch                  # mesh size bisection on log scale
    for _ in range(30):
        h = float(np.sqrt(lo * hi))
        ne = build(h).quads.shape[0]
        if abs(ne - target) < 0.05 * target:
            break
        lo, hi = (h, hi) if ne > target else (lo, h)
    mesh = build(h)
    f0 = 3100.0 / (2 * pitch * 1e-6)
    c = _stack(mesh, mats, 'xy', 2, (0.7 * f0, 1.3 * f0), points,
               info=dict(name=f'design_{seed}', seed=seed, pitch=pitch, backings=backs, coatings=coats,
                         piezo_um=piezo, mesh_um=h, target_elements=target))
    c.params = from_constants(k, euler)
    return c


def config5(n_cells=64, points=128):
    """The design set: ``n_cells`` cells with seeds 0..n_cells-1 (:func:`design_cell`)."""
    return [design_cell(s, points) for s in range(n_cells)]


# ------------------------------------------------------------------ Hex27 slabs
@dataclass
class HexCell:
    """A Hex27 slab cell: mesh, material records (piezo, electrode, backing layers from the top
    down; the PML is the last layer, or the piezo without layers), sweep band."""
    mesh: object
    materials: list
    band_hz: tuple
    points: int
    beta_dk: float = 0.0
    eta_eps: float = 0.0
    info: dict = field(default_factory=dict)
    params: np.ndarray = None

    ndisp = 3

    @property
    def aperture_m(self):
        return self.mesh.info['aperture_m']

    def frequencies(self, points=None):
        return np.linspace(self.band_hz[0], self.band_hz[1], points or self.points)

    def material_list(self, materials=None):
        """Reference-style material records (``hex_models.py:194-196``), PML elements excluded."""
        mats = materials if materials is not None else self.materials
        tm, tag_pml = self.mesh.info['tag_map'], self.mesh.info['tag_pml']
        return [dict(rho=mats[i].rho, C=mats[i].C, e=mats[i].e, eps=mats[i].eps,
                     indx=np.flatnonzero(self.mesh.hex_tag == tag)) for tag, i in tm.items() if tag != tag_pml]

    def pml_list(self, materials=None):
        """Bottom PML record: region 3 (z stretch) between the bottom and 2 pitch above it,
        dmax 1e11, n 3, with the PML region's own material (``hex_models.py:197-202``)."""
        mats = materials if materials is not None else self.materials
        info = self.mesh.info
        pm = mats[info['tag_map'][info['tag_pml']]]
        bottom = float(self.mesh.pos_m[self.mesh.sets['bottom'], 2].min())
        return [dict(region=3, indx=np.flatnonzero(self.mesh.hex_tag == info['tag_pml']),
                     za=bottom + info['pml_thickness_m'], zp=bottom, xa=0.0, xp=0.0, dmax=1e11, n=3,
                     rho=pm.rho, C=pm.C, e=pm.e, eps=pm.eps)]


def _cubic(c11, c12, c44, rho, eps_r):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    C[np.arange(3), np.arange(3)] = c11
    C[np.arange(3, 6), np.arange(3, 6)] = c44
    return mt.Material(rho, C, np.zeros((3, 6)), np.eye(3) * eps_r * mt.EPS0_LEGACY)


def hex_record_materials(euler, electrode, layers):
    """sp_baseline rotated by ``euler``, the electrode record and the layer records of the
    reference's slab templates (builtin legacy records, eps0 = 8.854e-12)."""
    C = np.array([[232966000000.0, 46890400000.0, 80234200000.0, -11026700000.0, 0, 0],
                  [46890400000.0, 232966000000.0, 80234200000.0, 11026700000.0, 0, 0],
                  [80234200000.0, 80234200000.0, 275364000000.0, 0, 0, 0],
                  [-11026700000.0, 11026700000.0, 0, 93899500000.0, 0, 0],
                  [0, 0, 0, 0, 93899500000.0, -11026700000.0],
                  [0, 0, 0, 0, -11026700000.0, 93038000000.0]])
    e = np.array([[0, 0, 0, 0, 2.59576, -1.58923], [-1.58923, 1.58923, 0, 2.59576, 0, 0],
                  [0.0821598, 0.0821598, 1.88197, 0, 0, 0]])
    eps = np.diag([3.6212859999999994e-10, 3.6212859999999994e-10, 3.833781999999999e-10])
    Cr, er, epsr = mt.rotate(euler, C, e, eps)
    G = 1e9
    aniso = np.array([[166, 64, 64, 0, 0, 0], [64, 195, 35, 0, 0, 0], [64, 35, 195, 0, 0, 0],
                      [0, 0, 0, 51, 0, 0], [0, 0, 0, 0, 80, 0], [0, 0, 0, 0, 0, 80]]) * G
    lib = dict(al=_al(), al_hex35=mt.isotropic(70e9, 0.35, 2700.0, 1.0, eps0=mt.EPS0_LEGACY),
               si_cubic_hex=_cubic(166 * G, 64 * G, 80 * G, 2330.0, 11.68),
               si_aniso_hex=mt.Material(2330.0, aniso, np.zeros((3, 6)), np.eye(3) * 11.8 * mt.EPS0_LEGACY),
               sio2=mt.isotropic(70e9, 0.17, 2200.0, 4.2, eps0=mt.EPS0_LEGACY),
               polysi=mt.isotropic(169e9, 0.22, 2320.0, 4.5, eps0=mt.EPS0_LEGACY),
               polysi_160=mt.isotropic(160e9, 0.22, 2320.0, 4.5, eps0=mt.EPS0_LEGACY))
    return [mt.Material(7450.0, Cr, er, epsr), lib[electrode]] + [lib[x] for x in layers]


# The reference's four slab templates (tests/data/sp_2p5d_*_layer.json configurations).
HEX_TEMPLATES = {
    'single': dict(pitch=1.085, aperture=0.217, electrode=0.17, substrate=6.51, mesh=0.217, layers=[],
                   electrode_mat='al', band=(1.5e9, 2.7e9)),
    'double': dict(pitch=1.085, aperture=0.217, electrode=0.17, substrate=0.6, mesh=0.217,
                   layers=[('si_cubic_hex', 6.51)], electrode_mat='al', band=(1.75e9, 2.0e9)),
    'triple': dict(pitch=1.085, aperture=0.217, electrode=0.17, substrate=0.6, mesh=0.217,
                   layers=[('sio2', 0.5), ('polysi', 6.51)], electrode_mat='al', band=(1.6e9, 2.0e9)),
    'quad': dict(pitch=1.0, aperture=0.25, electrode=0.17, substrate=0.6, mesh=0.2,
                 layers=[('sio2', 0.5), ('polysi_160', 1.0), ('si_aniso_hex', 6.0)], electrode_mat='al_hex35',
                 band=(1.8e9, 2.1e9)),
}


def hex_template(name, points=5, band=None):
    """The reference's sp_2p5d_<name>_layer slab (Euler (0, 48, 0), metallization 0.5)."""
    from .hexmesh import template_slab
    t = HEX_TEMPLATES[name]
    mesh = template_slab(name, t['pitch'], t['aperture'], t['electrode'], t['substrate'], 0.5, t['mesh'],
                         [th for _, th in t['layers']])
    mats = hex_record_materials((0, 48, 0), t['electrode_mat'], [m for m, _ in t['layers']])
    return HexCell(mesh=mesh, materials=mats, band_hz=band or t['band'], points=points,
                   info=dict(name=f'sp_2p5d_{name}_layer'))


# config 4: fR = 1.8278 GHz from a 201-point pilot over 1.0-3.0 GHz (pilot_resonance on the
# default slab), band fR * [0.92, 1.08]
BAND4 = (1.6816e9, 1.9741e9)


def config4(seed=0, mesh_um=0.047, points=201, band=BAND4, euler=(0, 52, 0), beta_dk=0.0, eta_eps=0.0):
    """Hex27 slab, one element across a 0.217 µm aperture: pitch 1.085 µm, Al electrodes
    0.08 pitch thick (metallization 0.5) on a 0.6 µm 3m film (sp_baseline constants with seeded
    ±5 % jitter; Euler (0, 52°, 0) is config 2's (0, -38°, 0) cut in slab axes, z = depth) over
    0.5 µm SiO2, 0.3 µm poly-Si (E = 169 GPa) and Si down to 5 pitch, plus a 2 pitch PML in Si.
    Element layers of about ``mesh_um`` everywhere (≈ 8k Hex27 elements at 0.047 µm)."""
    from .hexmesh import slab_cell, gmsh_segments
    pitch = 1.085
    nz = lambda h: max(1, int(np.ceil(h / mesh_um - 1e-12)))
    h_si = 5 * pitch - 0.6 - 0.5 - 0.3
    layers = [(0.5, nz(0.5)), (0.3, nz(0.3)), (h_si, nz(h_si))]
    d = 0.08 * pitch
    mesh = slab_cell(pitch, 0.217, d, 0.5, layers, 2 * pitch, nz(2 * pitch), nz(0.6), 0.6, nz(d), mesh_um)
    k = jittered_constants(seed)
    pz = piezo_material(k, euler, eta_eps, plane=None)
    mats = [pz, _al(), mt.isotropic(72e9, 0.17, 2200.0, 4.2), mt.isotropic(169e9, 0.22, 2320.0, 4.5),
            mt.isotropic(130e9, 0.28, 2330.0, 11.7)]
    c = HexCell(mesh=mesh, materials=mats, band_hz=band, points=points, beta_dk=beta_dk, eta_eps=eta_eps,
                info=dict(name='config4', constants=k, euler=euler, seed=seed, mesh_um=mesh_um))
    c.params = np.append(from_constants(k, euler, beta_dk, eta_eps), 7450.0)
    return c


WORKLOADS = dict(config1=config1, config2=config2, config3=config3, double_layer=double_layer, quad_layer=quad_layer,
                 config4=config4)


def assemble_cell(cell: Cell, materials=None):
    """K, M, D on the GPU for a workload cell (D = K_uu of the piezo, the damping operator)."""
    from .piezo_fem import assemble_device
    mats = materials if materials is not None else cell.materials
    nd = cell.ndisp + 1
    if nd == 3:
        mats = [mt.lift_projected(m) for m in mats]
    mask = np.zeros(len(mats), dtype=bool)
    mask[0] = True
    return assemble_device(cell.model, mats, cell.mat_idx, cell.pml, 2 * np.pi * cell.band_hz[1], nd,
                           damping_mask=mask)


def assemble_hex_cell(cell: HexCell, materials=None):
    """K, M, D on the GPU for a slab cell (D = K_uu of the piezo record)."""
    from . import hex27
    mesh = cell.mesh
    recs = cell.material_list(materials)
    return hex27.assemble_h27(mesh.hexas, mesh.pos_m, recs, cell.pml_list(materials), 2 * np.pi * cell.band_hz[1],
                              damping=[True] + [False] * (len(recs) - 1))


def build_hex_engine(cell: HexCell, gradient=False, **kw):
    """GPU FrequencySolver3D for a slab cell, with damping and (optionally) the dY11/dp context."""
    from . import hex27
    mesh = cell.mesh
    K, M, D = assemble_hex_cell(cell)
    bnd = hex27.boundary_idx_vect_3ds(hex27.find_saw_boundary_3ds(mesh), mesh.pos_m, 3 * mesh.N)
    eng = hex27.FrequencySolver3D(K, M, bnd, **kw)
    eng.set_damping(D, cell.beta_dk)
    if gradient:
        if cell.params is None:
            raise ValueError('this cell has no design parameters')
        eng.enable_gradient(mesh.hexas, mesh.pos_m, cell.material_list(), cell.pml_list(), 2 * np.pi * cell.band_hz[1],
                            cell.params)
    return eng


def hex_materials_for_params(cell: HexCell, p):
    """The slab cell's records with the piezo rebuilt from the 18-parameter vector ``p``."""
    from .params import NAMES
    k = {name: float(p[i]) for i, name in enumerate(NAMES[:12])}
    pz = piezo_material(k, tuple(float(a) for a in p[12:15]), float(p[16]), plane=None, rho=float(p[17]))
    return [pz] + list(cell.materials[1:])


def build_engine(cell: Cell, gradient=False, **kw):
    """GPU FrequencySolver for a cell, with damping and (optionally) the dY11/dp context."""
    from .sweep import FrequencySolver, bc_periodic
    if isinstance(cell, HexCell):
        return build_hex_engine(cell, gradient=gradient, **kw)
    K, M, D = assemble_cell(cell)
    eng = FrequencySolver(K, M, *bc_periodic(cell.model, cell.mesh.lines, ndisp=cell.ndisp), **kw)
    eng.set_damping(D, cell.beta_dk)
    if gradient:
        if cell.params is None:
            raise ValueError('this cell has no design parameters')
        eng.enable_gradient(cell.model, cell.mat_idx, cell.pml, 2 * np.pi * cell.band_hz[1], cell.params, 'xy')
    return eng


def pilot_resonance(builder, points=201, band=(1.0e9, 3.0e9), **kw):
    """fR of the main resonance of a cell from a pilot sweep (GPU, mixed precision refined to
    FP64 accuracy), and the sweep band fR * [0.92, 1.08]."""
    from .metrics import resonances
    cell = builder(points=points, band=band, **kw)
    eng = build_engine(cell)
    f = cell.frequencies()
    Q, _ = eng.sweep_charges(f, precision='mixed')
    fr = resonances(f, -1j * 2 * np.pi * f * Q)['fr_hz']
    return fr, (None if fr is None else (0.92 * fr, 1.08 * fr))


def materials_for_params(cell: Cell, p):
    """The cell's material list with the piezo rebuilt from parameter vector ``p``."""
    from .params import NAMES
    k = {name: float(p[i]) for i, name in enumerate(NAMES[:12])}
    pz = piezo_material(k, tuple(float(a) for a in p[12:15]), float(p[16]))
    return [pz] + list(cell.materials[1:])
