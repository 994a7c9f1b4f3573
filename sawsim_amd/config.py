"""Model fields: the reference's ``SimulationConfig`` (``src/sawsim/config.py:16-124``) with the
per-template defaults and ranges of ``MODEL_SPECS`` (``src/sawsim/sp_specs.py:4-80``), for the
templates this backend meshes itself, and the cell each configuration describes.

Templates and what they take:

* ``sp_double_layer``, ``sp_quad_layer``: the reference's structured 2D templates (ME0 or ME1),
  one / three backing layers of fixed roles;
* ``sp_stack``: the generic structured 2D stack, 0-6 backing layers and 0-3 coatings (the first
  coating measured from the piezo surface and embedding the electrodes);
* ``sp_2p5d_{single,double,triple,quad}_layer``: the Hex27 slabs (ME1, lossless, aperture
  required).

Every field of ``SimulationConfig`` is accepted and checked as the reference checks it (missing
fields take the template defaults first; values must be finite; ranges per template; stop above
start after the defaults are merged; layer counts and thicknesses; losses on 2D templates only).
Materials come from the builtin records this backend carries (``MATERIALS``); the substrate is the
``sp_baseline`` 3m record, rotated by the ZXZ Euler angles.  ``sp_single_layer`` and ``sp_tcsaw``
are refused: their reference meshes are unstructured Gmsh meshes that this backend does not
rebuild.
"""
from __future__ import annotations

import copy
import math
import numbers

import numpy as np

from . import materials as mt

# defaults of every template (sp_specs.py:4-7)
_BASE = dict(pitch_um=1.085, electrode_um=0.17, substrate_um=8.0, metal_ratio=0.5, mesh_um=0.5, start_ghz=1.5,
             stop_ghz=2.7, points=41, voltage=1.0, substrate_material='sp_baseline', electrode_material='al',
             euler_convention='zxz', euler_phi_deg=0.0, euler_theta_deg=-42.0, euler_psi_deg=0.0, mode_extension=1,
             layers=[], coatings=[], material_snapshots={}, aperture_um=None, beta_dk=0.0, eta_eps=0.0)
FIELDS = tuple(_BASE) + ('model_id',)


def _ranges(piezo_min, piezo_max, mesh_min=0.25, mesh_max=0.8):
    """Numeric field ranges of the fixed-layer templates (sp_specs.py:9-14)."""
    return dict(pitch_um=(0.8, 1.5), electrode_um=(0.08, 0.3), substrate_um=(piezo_min, piezo_max),
                metal_ratio=(0.3, 0.7), mesh_um=(mesh_min, mesh_max), start_ghz=(0.5, 5), stop_ghz=(0.5, 5),
                points=(3, 1201), voltage=(0.01, 10))


def _layers(*pairs):
    return [dict(material_id=m, thickness_um=h) for m, h in pairs]


def _spec(layers, substrate_um, ranges, **extra):
    d = copy.deepcopy(_BASE)
    d.update(substrate_um=substrate_um, layers=_layers(*layers))
    s = dict(defaults=d, ranges=ranges, layer_count=(len(layers), len(layers)), coating_count=(0, 0),
             layer_thickness=(0.02, 15.0), mode_extensions=(0, 1))
    s.update(extra)
    return s


SPECS = {
    'sp_double_layer': _spec([('si_isotropic', 8.33)], 0.6, _ranges(0.1, 3)),
    'sp_quad_layer': _spec([('sio2', 0.5), ('polysi', 1.0), ('si_isotropic', 7.83)], 0.6, _ranges(0.1, 3)),
    'sp_stack': _spec([], 8.0, dict(pitch_um=(0.1, 20), electrode_um=(0.005, 2), substrate_um=(0.05, 100),
                                    metal_ratio=(0.15, 0.85), mesh_um=(0.01, 1), start_ghz=(0.02, 20),
                                    stop_ghz=(0.02, 20), points=(3, 1201), voltage=(0.01, 10)),
                      layer_count=(0, 6), coating_count=(0, 3), layer_thickness=(0.005, 100.0)),
}
for _name, _piezo, _lay, _band in [('single', 6.51, [], (1.5, 2.7)),
                                    ('double', 0.6, [('si_cubic_hex', 6.51)], (1.75, 2.0)),
                                    ('triple', 0.6, [('sio2', 0.5), ('polysi', 6.51)], (1.6, 2.0)),
                                    ('quad', 0.6, [('sio2', 0.5), ('polysi_160', 1.0), ('si_aniso_hex', 6.0)],
                                     (1.8, 2.1))]:
    _r = _ranges(0.1 if _lay else 4, 3 if _lay else 12, 0.1, 0.5)
    _r['aperture_um'] = (0.05, 0.5)
    _s = _spec(_lay, _piezo, _r, mode_extensions=(1,), hex=_name)
    _s['defaults'].update(start_ghz=_band[0], stop_ghz=_band[1], euler_theta_deg=48.0,
                          mesh_um=0.2 if _name == 'quad' else 0.217, aperture_um=0.25 if _name == 'quad' else 0.217)
    if _name == 'quad':
        _s['defaults'].update(pitch_um=1.0, electrode_material='al_hex35')
    SPECS[f'sp_2p5d_{_name}_layer'] = _s

UNSUPPORTED = {'sp_single_layer': 'its reference mesh is an unstructured Gmsh mesh (Mesh.Algorithm 8)',
               'sp_tcsaw': 'its reference mesh is an unstructured Gmsh mesh (Mesh.Algorithm 8)'}


def _cubic(c11, c12, c44, rho, eps_r):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    C[[0, 1, 2], [0, 1, 2]] = c11
    C[[3, 4, 5], [3, 4, 5]] = c44
    return mt.Material(rho, C, np.zeros((3, 6)), np.eye(3) * eps_r * mt.EPS0_LEGACY)


def _aniso_si():
    G = 1e9
    C = np.array([[166, 64, 64, 0, 0, 0], [64, 195, 35, 0, 0, 0], [64, 35, 195, 0, 0, 0],
                  [0, 0, 0, 51, 0, 0], [0, 0, 0, 0, 80, 0], [0, 0, 0, 0, 0, 80]]) * G
    return mt.Material(2330.0, C, np.zeros((3, 6)), np.eye(3) * 11.8 * mt.EPS0_LEGACY)


# builtin records this backend carries: id -> (roles, builder); the legacy records use
# eps0 = 8.854e-12 as the reference's materials.py:15 does
MATERIALS = {
    'al': (('electrode',), lambda: mt.isotropic(70e9, 0.33, 2700.0, 1.0, eps0=mt.EPS0_LEGACY)),
    'al_hex35': (('electrode',), lambda: mt.isotropic(70e9, 0.35, 2700.0, 1.0, eps0=mt.EPS0_LEGACY)),
    'si_isotropic': (('layer',), lambda: mt.isotropic(170e9, 0.28, 2329.0, 11.7, eps0=mt.EPS0_LEGACY)),
    'sio2': (('layer',), lambda: mt.isotropic(70e9, 0.17, 2200.0, 4.2, eps0=mt.EPS0_LEGACY)),
    'polysi': (('layer',), lambda: mt.isotropic(169e9, 0.22, 2320.0, 4.5, eps0=mt.EPS0_LEGACY)),
    'polysi_160': (('layer',), lambda: mt.isotropic(160e9, 0.22, 2320.0, 4.5, eps0=mt.EPS0_LEGACY)),
    'si_cubic_hex': (('layer',), lambda: _cubic(166e9, 64e9, 80e9, 2330.0, 11.68)),
    'si_aniso_hex': (('layer',), _aniso_si),
    'sp_baseline': (('substrate',), None),
}


def sp_baseline(euler_deg):
    """The sp_baseline 3m record (the reference's builtin substrate), rotated ZXZ by euler_deg."""
    C = np.array([[232966000000.0, 46890400000.0, 80234200000.0, -11026700000.0, 0, 0],
                  [46890400000.0, 232966000000.0, 80234200000.0, 11026700000.0, 0, 0],
                  [80234200000.0, 80234200000.0, 275364000000.0, 0, 0, 0],
                  [-11026700000.0, 11026700000.0, 0, 93899500000.0, 0, 0],
                  [0, 0, 0, 0, 93899500000.0, -11026700000.0],
                  [0, 0, 0, 0, -11026700000.0, 93038000000.0]])
    e = np.array([[0, 0, 0, 0, 2.59576, -1.58923], [-1.58923, 1.58923, 0, 2.59576, 0, 0],
                  [0.0821598, 0.0821598, 1.88197, 0, 0, 0]])
    eps = np.diag([3.6212859999999994e-10, 3.6212859999999994e-10, 3.833781999999999e-10])
    return mt.Material(7450.0, *mt.rotate(tuple(euler_deg), C, e, eps))


def finite_number(v):
    return isinstance(v, numbers.Real) and not isinstance(v, bool) and math.isfinite(v)


def integer(v):
    return isinstance(v, numbers.Integral) and not isinstance(v, bool)


def _layer(d, where):
    if not isinstance(d, dict):
        raise ValueError(f'{where}: a layer is a dict with material_id and thickness_um')
    extra = set(d) - {'material_id', 'thickness_um', 'euler_phi_deg', 'euler_theta_deg', 'euler_psi_deg'}
    if extra:
        raise ValueError(f'{where}: unknown layer fields {sorted(extra)}')
    out = dict(material_id=d.get('material_id'), thickness_um=d.get('thickness_um'), euler_phi_deg=0.0,
               euler_theta_deg=0.0, euler_psi_deg=0.0)
    out.update({k: v for k, v in d.items() if k.startswith('euler_')})
    if not isinstance(out['material_id'], str):
        raise ValueError(f'{where}: material_id must be a string')
    if not (finite_number(out['thickness_um']) and 0 < out['thickness_um'] <= 30):
        raise ValueError(f"{where}: thickness_um must be in (0, 30], not {out['thickness_um']!r}")
    for k in ('euler_phi_deg', 'euler_theta_deg', 'euler_psi_deg'):
        if not (finite_number(out[k]) and -360 <= out[k] <= 360):
            raise ValueError(f'{where}: {k} must be a finite angle in [-360, 360]')
    return out


def validate(model_id, fields: dict) -> dict:
    """The full, checked configuration of ``model_id`` with ``fields`` over its defaults
    (``config.py:49-124``); raises ValueError on the first violation."""
    if model_id in UNSUPPORTED:
        raise ValueError(f'{model_id} is not built by this backend: {UNSUPPORTED[model_id]}')
    if model_id not in SPECS:
        raise ValueError(f'unknown model_id {model_id!r}')
    unknown = set(fields) - set(FIELDS)
    if unknown:
        raise ValueError(f'unknown fields {sorted(unknown)}')
    spec = SPECS[model_id]
    c = copy.deepcopy(spec['defaults'])
    c.update(copy.deepcopy(fields))
    c['model_id'] = model_id
    if c['euler_convention'] in ('reference_zxz', 'zxz_intrinsic'):
        c['euler_convention'] = 'zxz'
    if c['euler_convention'] != 'zxz':
        raise ValueError("euler_convention must be 'zxz'")
    if not (integer(c['mode_extension']) and c['mode_extension'] in (0, 1)):
        raise ValueError('mode_extension must be the integer 0 or 1')
    if not integer(c['points']):
        raise ValueError(f"points must be an integer, not {c['points']!r}")
    for k in ('euler_phi_deg', 'euler_theta_deg', 'euler_psi_deg'):
        if not (finite_number(c[k]) and -360 <= c[k] <= 360):
            raise ValueError(f'{k} must be a finite angle in [-360, 360], not {c[k]!r}')
    positive = ('pitch_um', 'electrode_um', 'substrate_um', 'mesh_um', 'start_ghz', 'stop_ghz', 'voltage')
    for k in positive:
        if not (finite_number(c[k]) and c[k] > 0):
            raise ValueError(f'{k} must be a finite number > 0, not {c[k]!r}')
    if not (finite_number(c['metal_ratio']) and 0 < c['metal_ratio'] < 1):
        raise ValueError(f"metal_ratio must be in (0, 1), not {c['metal_ratio']!r}")
    if not (finite_number(c['beta_dk']) and 0 <= c['beta_dk'] <= 1e-10):
        raise ValueError(f"beta_dk must be in [0, 1e-10] s, not {c['beta_dk']!r}")
    if not (finite_number(c['eta_eps']) and 0 <= c['eta_eps'] <= 0.1):
        raise ValueError(f"eta_eps must be in [0, 0.1], not {c['eta_eps']!r}")
    if c['aperture_um'] is not None and not (finite_number(c['aperture_um']) and c['aperture_um'] > 0):
        raise ValueError(f"aperture_um must be a finite number > 0, not {c['aperture_um']!r}")
    if not isinstance(c['material_snapshots'], dict) or c['material_snapshots']:
        raise ValueError('material_snapshots are not supported: this backend uses its builtin records')
    for key in ('layers', 'coatings'):
        if not isinstance(c[key], list):
            raise ValueError(f'{key} must be a list')
        c[key] = [_layer(d, f'{key}[{i}]') for i, d in enumerate(c[key])]
    # materials: known records in their roles
    refs = [(c['substrate_material'], 'substrate'), (c['electrode_material'], 'electrode')]
    refs += [(d['material_id'], 'layer') for d in c['layers'] + c['coatings']]
    for mid, role in refs:
        if mid not in MATERIALS:
            raise ValueError(f'unknown material {mid!r} (this backend carries {sorted(MATERIALS)})')
        if role not in MATERIALS[mid][0]:
            raise ValueError(f'material {mid!r} cannot be used as {role}')
    hex_ = 'hex' in spec
    if hex_ and c['aperture_um'] is None:
        raise ValueError('the 2.5D slab templates need aperture_um')
    if not hex_ and c['aperture_um'] is not None:
        raise ValueError('aperture_um applies to the 2.5D Hex27 slab templates only')
    if hex_ and (c['beta_dk'] or c['eta_eps']):
        raise ValueError('material loss (beta_dk / eta_eps) is supported on the 2D templates only')
    if c['mode_extension'] not in spec['mode_extensions']:
        raise ValueError('this model does not support the chosen mode_extension')
    lo, hi = spec['layer_count']
    if not lo <= len(c['layers']) <= hi:
        raise ValueError(f'this model needs {lo} backing layers' if lo == hi else
                         f'this model takes {lo}-{hi} backing layers')
    lo, hi = spec['coating_count']
    if not lo <= len(c['coatings']) <= hi:
        raise ValueError('this model takes no coatings' if hi == 0 else f'this model takes {lo}-{hi} coatings')
    if c['coatings'] and c['coatings'][0]['thickness_um'] <= c['electrode_um']:
        raise ValueError('the first coating is measured from the piezo surface and embeds the electrodes: '
                         'it must be thicker than the electrodes')
    lo, hi = spec['layer_thickness']
    for d in c['layers'] + c['coatings']:
        if not lo <= d['thickness_um'] <= hi:
            raise ValueError(f'layer thickness outside [{lo}, {hi}] um')
    for k, (lo, hi) in spec['ranges'].items():
        if not lo <= c[k] <= hi:
            raise ValueError(f'{k} outside [{lo}, {hi}]: {c[k]!r}')
    if c['stop_ghz'] <= c['start_ghz']:
        raise ValueError('stop_ghz must be above start_ghz')
    return c


def _layer_material(d):
    m = MATERIALS[d['material_id']][1]()
    euler = (d['euler_phi_deg'], d['euler_theta_deg'], d['euler_psi_deg'])
    if any(euler):       # layers rotate only with non-zero angles (material_runtime.py:20-21)
        m = mt.Material(m.rho, *mt.rotate(euler, m.C, m.e, m.eps))
    return m


def build_cell(c: dict):
    """The cell (mesh, materials, PML, band) of a validated configuration ``c``."""
    from . import workloads as wl
    from .mesh import double_layer_cell, quad_layer_cell, stack_cell
    euler = (c['euler_phi_deg'], c['euler_theta_deg'], c['euler_psi_deg'])
    piezo = sp_baseline(euler)
    electrode = MATERIALS[c['electrode_material']][1]()
    layers = [_layer_material(d) for d in c['layers']]
    coatings = [_layer_material(d) for d in c['coatings']]
    band = (c['start_ghz'] * 1e9, c['stop_ghz'] * 1e9)
    th = [d['thickness_um'] for d in c['layers']]
    if 'hex' in SPECS[c['model_id']]:
        from .hexmesh import template_slab
        name = SPECS[c['model_id']]['hex']
        mesh = template_slab(name, c['pitch_um'], c['aperture_um'], c['electrode_um'], c['substrate_um'],
                             c['metal_ratio'], c['mesh_um'], th)
        return wl.HexCell(mesh=mesh, materials=[piezo, electrode] + layers, band_hz=band, points=c['points'],
                          info=dict(name=c['model_id']))
    mid = c['model_id']
    args = (c['pitch_um'], c['metal_ratio'], c['electrode_um'], c['substrate_um'])
    if mid == 'sp_double_layer':
        mesh = double_layer_cell(c['pitch_um'], c['electrode_um'], c['metal_ratio'], c['substrate_um'], th[0],
                                 c['mesh_um'])
    elif mid == 'sp_quad_layer':
        mesh = quad_layer_cell(c['pitch_um'], c['metal_ratio'], c['electrode_um'], c['substrate_um'], th[0], th[1],
                               th[2], c['mesh_um'])
    else:
        mesh = stack_cell(*args, th, [d['thickness_um'] for d in c['coatings']], c['mesh_um'], pml_pitches=2.0)
    me0 = c['mode_extension'] == 0
    if me0:
        piezo = mt.project(piezo, 'xy')
    if c['eta_eps']:          # dielectric loss on the piezo only (models.py:68-70)
        piezo = mt.Material(piezo.rho, piezo.C, piezo.e, piezo.eps * (1 - 1j * c['eta_eps']))
    cell = wl._stack(mesh, [piezo, electrode] + layers + coatings, 'xy' if me0 else None, 2 if me0 else 3, band,
                     c['points'], c['beta_dk'], c['eta_eps'], info=dict(name=mid))
    return cell
