"""Model -> sweep -> Result, with the reference's result-directory layout.

Mirrors ``sawsim.Model`` / ``sawsim.sweep`` / ``sawsim.Result`` (``src/sawsim/api.py:12-90``) and
the artifacts ``run_simulation`` writes (``src/sawsim/solver.py:56-157``) for the cells this
backend builds itself (the structured SP templates and the synthetic stack cells of
``workloads``):

* ``input.json`` (the model's fields), ``progress.json`` (stage events),
* ``mesh.npz`` {position_m, quads, material_tags},
* ``curve.json`` {frequency_ghz, real, imag, magnitude}, ``admittance.npz`` {frequency_hz,
  charge, admittance}, ``admittance.csv`` (frequency_hz, real_y, imag_y, abs_y),
* ``fields.npz``: the FP64 re-solve at the largest |Y| sample, displacement scaled by SCALE
  (``solver.py:131-140``),
* ``gradient.npz`` {frequency_hz, names, dy11_dp} when the gradient is requested (no
  counterpart in the reference),
* ``metadata.json`` and ``manifest.json`` (artifact list).

The admittance convention is the reference's passive sign, Y = -i w Q / V, for the plane
cells.  The Hex27 slab templates follow ``run_hex_simulation`` (``src/sawsim/hex_solver.py:88-120``):
Y = +i w Q / V, and charge and admittance per unit aperture (raw slice values kept as
``raw_charge_c``/``raw_admittance_s``).  Images, the material snapshot and Gmsh meshes are not
produced.
"""
from __future__ import annotations

import json
import shutil
import tempfile
import time
from pathlib import Path
from typing import Callable, Optional

import numpy as np

from . import config as _cfg

TEMPLATES = {
    'sp_double_layer': 'structured double-layer SP cell (reference template; ME0 or ME1)',
    'sp_quad_layer': 'structured quad-layer SP cell (reference template; ME0 or ME1)',
    'sp_stack': 'generic structured SP stack, 0-6 backing layers and 0-3 coatings (reference template)',
    'sp_2p5d_single_layer': 'Hex27 slab, piezo on its own PML (reference template)',
    'sp_2p5d_double_layer': 'Hex27 slab, piezo on cubic Si (reference template)',
    'sp_2p5d_triple_layer': 'Hex27 slab, piezo / SiO2 / poly-Si (reference template)',
    'sp_2p5d_quad_layer': 'Hex27 slab, piezo / SiO2 / poly-Si / anisotropic Si (reference template)',
    'config1': 'sp_stack-style 3m film on Si, plane strain (ME0), 4312 elements',
    'config2': 'config1 with a SiO2 overlay, refined mesh (ME0), 16848 elements',
    'config3': 'config1 with Rayleigh damping and dielectric loss (ME0)',
    'config4': 'Hex27 slab, 3m film / SiO2 / poly-Si / Si, ~8k elements',
}
# the synthetic workload cells take the sweep fields only; their default bands and point counts
_SYNTH = {'config1': ((1.3143, 1.5429), 101), 'config2': ((2.3114, 2.7134), 401),
          'config3': ((1.3143, 1.5429), 201), 'config4': ((1.6816, 1.9741), 201)}
_SYNTH_FIELDS = dict(points=None, start_ghz=None, stop_ghz=None, voltage=1.0)
# this backend's own fields, beside the reference's
_BACKEND = dict(precision='mixed', gradient=False, batch=None)


def _check_sweep_fields(c: dict):
    """The reference's bounds on the sweep fields of a synthetic cell (``config.py:36-39,
    122-123``), after the cell's default band and point count are filled in: an integer
    3 <= points <= 1201, finite start and stop > 0 with stop > start, finite voltage > 0."""
    if not (_cfg.integer(c['points']) and 3 <= c['points'] <= 1201):
        raise ValueError(f"points must be an integer in [3, 1201], not {c['points']!r}")
    for k in ('start_ghz', 'stop_ghz', 'voltage'):
        if not (_cfg.finite_number(c[k]) and c[k] > 0):
            raise ValueError(f'{k} must be a finite number > 0, not {c[k]!r}')
    if c['stop_ghz'] <= c['start_ghz']:
        raise ValueError('stop_ghz must be above start_ghz')


def _check_backend(c: dict):
    """precision 'mixed' or 'fp64', gradient a bool, batch a positive integer or None."""
    if c['precision'] not in ('mixed', 'fp64'):
        raise ValueError(f"precision must be 'mixed' or 'fp64', not {c['precision']!r}")
    if not isinstance(c['gradient'], bool):
        raise ValueError(f"gradient must be True or False, not {c['gradient']!r}")
    if c['batch'] is not None and not (_cfg.integer(c['batch']) and c['batch'] >= 1):
        raise ValueError(f"batch must be a positive integer, not {c['batch']!r}")


def _atomic_json(path: Path, data):
    tmp = path.with_suffix('.tmp')
    tmp.write_text(json.dumps(data, allow_nan=False), encoding='utf-8')
    tmp.replace(path)


class Model:
    """A unit-cell model: template id plus fields (``sawsim.Model``, ``api.py:12-35``).

    Reference templates (``sp_double_layer``, ``sp_quad_layer``, ``sp_stack``, ``sp_2p5d_*``) take
    every ``SimulationConfig`` field (pitch_um, electrode_um, substrate_um, metal_ratio, mesh_um,
    layers, coatings, Euler angles, mode_extension, aperture_um, beta_dk, eta_eps, start_ghz,
    stop_ghz, points, voltage, ...), over the template's defaults and checked against its ranges
    (:mod:`sawsim_amd.config`).  The synthetic workload cells (``config1`` .. ``config4``) take the
    sweep fields only.  Every model also takes this backend's precision ('mixed' | 'fp64'),
    gradient and batch.  Invalid fields raise ValueError, as the reference's validation does.
    """

    def __init__(self, model_id: str = 'sp_double_layer', **params):
        backend = {k: params.pop(k) for k in list(params) if k in _BACKEND}
        if model_id in _SYNTH:
            unknown = set(params) - set(_SYNTH_FIELDS)
            if unknown:
                raise ValueError(f'unknown fields {sorted(unknown)} (a synthetic workload cell takes '
                                 f'{sorted(_SYNTH_FIELDS)})')
            (lo, hi), pts = _SYNTH[model_id]
            conf = dict(_SYNTH_FIELDS, **params)
            conf['points'] = pts if conf['points'] is None else conf['points']
            conf['start_ghz'] = lo if conf['start_ghz'] is None else conf['start_ghz']
            conf['stop_ghz'] = hi if conf['stop_ghz'] is None else conf['stop_ghz']
            _check_sweep_fields(conf)
        elif model_id in TEMPLATES or model_id in _cfg.UNSUPPORTED:
            conf = _cfg.validate(model_id, params)
        else:
            raise ValueError(f'unknown model_id {model_id!r}; available: {sorted(TEMPLATES)}')
        self.model_id = model_id
        self.config = dict(conf, **dict(_BACKEND, **backend), model_id=model_id)
        _check_backend(self.config)
        if self.config['gradient'] and model_id not in _SYNTH:
            raise ValueError('dY11/dp is computed for the parameterised cells config1-config4 only')

    @classmethod
    def from_json(cls, path) -> 'Model':
        data = json.loads(Path(path).read_text(encoding='utf-8'))
        return cls(**data)

    @staticmethod
    def templates() -> dict:
        return dict(TEMPLATES)

    def to_dict(self) -> dict:
        return dict(self.config)

    def cell(self):
        """The workload cell (mesh, materials, band) this model sweeps."""
        from . import workloads
        c = self.config
        if self.model_id not in _SYNTH:
            return _cfg.build_cell(c)
        band = (c['start_ghz'] * 1e9, c['stop_ghz'] * 1e9)
        return workloads.WORKLOADS[self.model_id](points=int(c['points']), band=band)

    def __repr__(self):
        c = self.config
        return f"Model({self.model_id!r}, points={c['points']}, {c['start_ghz']}-{c['stop_ghz']} GHz)"


class Result:
    """Results of one sweep, backed by an output directory."""

    def __init__(self, output_dir):
        self.dir = Path(output_dir)
        self.manifest = self._load('manifest.json')
        self.metadata = self._load('metadata.json')
        self.curve = self._load('curve.json')

    def _load(self, name):
        p = self.dir / name
        return json.loads(p.read_text(encoding='utf-8')) if p.exists() else {}

    @property
    def frequency_ghz(self) -> np.ndarray:
        return np.asarray(self.curve.get('frequency_ghz', []), dtype=float)

    @property
    def admittance(self) -> np.ndarray:
        return (np.asarray(self.curve.get('real', []), dtype=float)
                + 1j * np.asarray(self.curve.get('imag', []), dtype=float))

    @property
    def magnitude(self) -> np.ndarray:
        return np.abs(self.admittance)

    @property
    def peak_frequency_ghz(self) -> Optional[float]:
        return self.metadata.get('peak_frequency_ghz')

    def resonances(self) -> dict:
        from .metrics import resonances
        return resonances(self.frequency_ghz * 1e9, self.admittance)

    def artifact(self, name: str) -> Path:
        return self.dir / name

    def save(self, target) -> Path:
        target = Path(target)
        if target.resolve() != self.dir.resolve():
            shutil.copytree(self.dir, target, dirs_exist_ok=True)
        return target

    def __repr__(self):
        return f'Result({self.dir}, {len(self.frequency_ghz)} points, peak={self.peak_frequency_ghz} GHz)'


def write_sweep_results(out: Path, fre, charge, voltage, *, mesh=None, fields=None, gradient=None, metadata=None):
    """Write the curve, admittance, mesh, fields, gradient, metadata and manifest artifacts.

    ``fre`` Hz, ``charge`` complex Q per frequency; ``mesh`` = (position_m, quads, tags);
    ``fields`` = dict for fields.npz; ``gradient`` = (names, dY (P, k)).
    """
    out = Path(out)
    fre = np.asarray(fre, dtype=float)
    q = np.asarray(charge, dtype=complex)
    y = -1j * 2 * np.pi * fre * q / voltage
    _atomic_json(out / 'curve.json', dict(frequency_ghz=(fre * 1e-9).tolist(), real=y.real.tolist(),
                                          imag=y.imag.tolist(), magnitude=np.abs(y).tolist()))
    np.savez_compressed(out / 'admittance.npz', frequency_hz=fre, charge=q, admittance=y)
    np.savetxt(out / 'admittance.csv', np.column_stack([fre, y.real, y.imag, np.abs(y)]), delimiter=',',
               header='frequency_hz,real_y,imag_y,abs_y', comments='')
    if mesh is not None:
        pos, quads, tags = mesh
        np.savez_compressed(out / 'mesh.npz', position_m=pos, quads=quads, material_tags=tags)
    if fields is not None:
        np.savez_compressed(out / 'fields.npz', **fields)
    if gradient is not None:
        names, dy = gradient
        np.savez_compressed(out / 'gradient.npz', frequency_hz=fre, names=np.asarray(names), dy11_dp=np.asarray(dy))
    peak = int(np.argmax(np.abs(y)))
    meta = dict(frequency_count=int(fre.size), peak_frequency_ghz=float(fre[peak] * 1e-9),
                peak_is_boundary=peak in (0, fre.size - 1),
                admittance_convention='passive: Y = -i w Q_reaction / V',
                units={'frequency': 'Hz', 'position': 'm', 'displacement': 'm', 'potential': 'V',
                       'admittance': 'S/m (unit out-of-plane aperture)', 'charge': 'C/m'})
    meta.update(metadata or {})
    _atomic_json(out / 'metadata.json', meta)
    media = {'.json': 'application/json', '.csv': 'text/csv', '.npz': 'application/octet-stream'}
    artifacts = [dict(name=p.name, kind='data', media_type=media[p.suffix]) for p in sorted(out.iterdir())
                 if p.suffix in media and p.name not in ('manifest.json', 'progress.json')]
    artifacts.append(dict(name='manifest.json', kind='manifest', media_type='application/json'))
    _atomic_json(out / 'manifest.json', dict(model_id=meta.get('model_id'), artifacts=artifacts))
    return y


def _write_mesh_only(model, cell, out, started):
    """Mesh, metadata (task_kind 'mesh') and manifest of a mesh-only run (``solver.py:85-100``)."""
    from .workloads import HexCell
    out = Path(out)
    if isinstance(cell, HexCell):
        mesh = cell.mesh
        np.savez_compressed(out / 'mesh.npz', position_m=mesh.pos_m, hexahedra=mesh.hexas, material_tags=mesh.hex_tag)
        nodes, elements, dim, comps = int(mesh.pos_m.shape[0]), int(mesh.hexas.shape[0]), 3, 3
    else:
        m = cell.model
        np.savez_compressed(out / 'mesh.npz', position_m=m.pos_m, quads=m.quads, material_tags=m.elem_tag)
        nodes, elements, dim, comps = int(m.N), int(m.ne), 2, int(cell.ndisp)
    _atomic_json(out / 'metadata.json', dict(task_kind='mesh', model_id=model.model_id, nodes=nodes,
                                             elements=elements, mesh_dimension=dim, displacement_components=comps,
                                             elapsed_seconds=time.time() - started, units={'position': 'm'},
                                             input_file='input.json'))
    media = {'.json': 'application/json', '.npz': 'application/octet-stream'}
    artifacts = [dict(name=p.name, kind='data', media_type=media[p.suffix]) for p in sorted(out.iterdir())
                 if p.suffix in media and p.name not in ('manifest.json', 'progress.json')]
    artifacts.append(dict(name='manifest.json', kind='manifest', media_type='application/json'))
    _atomic_json(out / 'manifest.json', dict(model_id=model.model_id, artifacts=artifacts))


def _sweep_hex(model, cell, fre, out, progress, started):
    """Hex27 slab sweep and artifacts (``hex_solver.py:88-120``): Y = +i w Q / V per unit aperture."""
    from .hex27 import KM_SCALE
    from .workloads import build_hex_engine
    c = model.config
    eng = build_hex_engine(cell)
    progress('sweep', 0, fre.size)
    V = float(c['voltage'])
    Q, info = eng.sweep_charges(fre, V, precision=c['precision'], batch=c['batch'],
                                on_chunk=lambda done, total: progress('sweep', done, total))
    if not np.all(np.isfinite(Q)):
        raise ValueError('non-finite charge in the sweep')
    progress('postprocess', fre.size, fre.size)
    ap = cell.aperture_m
    raw_y = 1j * 2 * np.pi * fre * Q / V
    y = raw_y / ap
    _atomic_json(out / 'curve.json', dict(frequency_ghz=(fre * 1e-9).tolist(), real=y.real.tolist(),
                                          imag=y.imag.tolist(), magnitude=np.abs(y).tolist()))
    np.savez_compressed(out / 'admittance.npz', frequency_hz=fre, charge=Q / ap, admittance=y, raw_charge_c=Q,
                        raw_admittance_s=raw_y, aperture_m=ap)
    np.savetxt(out / 'admittance.csv', np.column_stack([fre, y.real, y.imag, np.abs(y), raw_y.real, raw_y.imag,
                                                        Q.real, Q.imag]), delimiter=',',
               header='frequency_hz,real_y_s_per_m,imag_y_s_per_m,abs_y_s_per_m,raw_real_y_s,raw_imag_y_s,'
                      'raw_real_reaction_q_c,raw_imag_reaction_q_c', comments='')
    mesh = cell.mesh
    quads, qtags = mesh.front_section()
    np.savez_compressed(out / 'mesh.npz', position_m=mesh.pos_m, hexas27=mesh.hexas, material_tags=mesh.hex_tag,
                        quads=quads, preview_material_tags=qtags, mesh_view_axes=[0, 2])
    peak = int(np.argmax(np.abs(y)))
    disp, _, _ = eng.solve(2 * np.pi * fre[peak], V)
    n = mesh.N
    np.savez_compressed(out / 'fields.npz', position_m=mesh.pos_m, displacement_m=disp[:3 * n].reshape(n, 3) * KM_SCALE,
                        potential_v=disp[3 * n:], frequency_hz=fre[peak], hexas27=mesh.hexas, quads=quads,
                        mesh_dimension=3, displacement_components=3, component_labels=['ux', 'uy', 'uz'],
                        mesh_view_axes=[0, 2])
    meta = dict(model_id=model.model_id, model='2.5D periodic thin slice / Hex27', mesh_dimension=3,
                frequency_count=int(fre.size), peak_frequency_ghz=float(fre[peak] * 1e-9),
                peak_is_boundary=peak in (0, fre.size - 1), aperture_m=ap, aperture_um=ap * 1e6,
                nodes=int(n), elements=int(mesh.ne), dofs=int(4 * n), active_dofs=int(eng.n_active),
                pml_reference_frequency_hz=float(cell.band_hz[1]),
                backend='MI355X batched multifrontal LU (HIP, libsawfem_hip)', precision=c['precision'],
                sign_convention='Y = +i w Q / V (reaction charge of the potential equation)',
                normalization='charge = raw_charge_c / aperture_m; admittance = raw_admittance_s / aperture_m',
                units={'frequency': 'Hz', 'position': 'm', 'displacement': 'm', 'potential': 'V',
                       'admittance': 'S/m', 'raw_admittance_s': 'S', 'raw_charge_c': 'C reaction'},
                refinement=[{k: v for k, v in it.items() if k != 'dq_hist'} for it in info],
                elapsed_seconds=time.time() - started, field_selection='maximum sampled admittance magnitude')
    _atomic_json(out / 'metadata.json', meta)
    media = {'.json': 'application/json', '.csv': 'text/csv', '.npz': 'application/octet-stream'}
    artifacts = [dict(name=p.name, kind='data', media_type=media[p.suffix]) for p in sorted(out.iterdir())
                 if p.suffix in media and p.name not in ('manifest.json', 'progress.json')]
    artifacts.append(dict(name='manifest.json', kind='manifest', media_type='application/json'))
    _atomic_json(out / 'manifest.json', dict(model_id=model.model_id, artifacts=artifacts))


def sweep(model: Model | dict, output_dir=None, *, mesh_only: bool = False, on_progress: Optional[Callable] = None,
          should_cancel: Optional[Callable] = None) -> Result:
    """Run the frequency sweep of ``model`` on the GPU and return a Result.

    A fresh temporary directory is used when ``output_dir`` is None; a directory that already
    holds results is refused (``FileExistsError``), as in the reference.  ``mesh_only`` writes
    the mesh, metadata and manifest and stops before any solve (``solver.py:85-100``).
    Cancellation, as in ``run_simulation`` (``solver.py:64-66``): a file named ``cancel`` in
    the output directory or ``should_cancel()`` returning true, checked at every progress event
    (stage boundaries and after each frequency chunk), raises ``InterruptedError``.
    """
    from . import params as prm
    from .piezo_fem import SCALE
    from .workloads import build_engine
    model = model if isinstance(model, Model) else Model(**model)
    c = model.config
    out = Path(output_dir) if output_dir else Path(tempfile.mkdtemp(prefix='sawsim_amd_'))
    out.mkdir(parents=True, exist_ok=True)
    if any((out / n).exists() for n in ('admittance.npz', 'manifest.json', 'mesh.npz')):
        raise FileExistsError(f'{out} already holds results')
    _atomic_json(out / 'input.json', model.to_dict())
    started = time.time()

    def check_cancel():
        if (out / 'cancel').exists() or (should_cancel is not None and should_cancel()):
            raise InterruptedError('sweep cancelled')

    def progress(stage, done, total):
        check_cancel()
        event = dict(stage=stage, completed=done, total=total, elapsed_seconds=round(time.time() - started, 2))
        _atomic_json(out / 'progress.json', event)
        if on_progress is not None:
            on_progress(event)
        check_cancel()

    cell = model.cell()
    fre = cell.frequencies()
    if mesh_only:
        progress('mesh', 0, 1)
        _write_mesh_only(model, cell, out, started)
        progress('complete', 1, 1)
        return Result(out)
    progress('assembly', 0, fre.size)
    from .workloads import HexCell
    if isinstance(cell, HexCell):
        _sweep_hex(model, cell, fre, out, progress, started)
        progress('complete', fre.size, fre.size)
        return Result(out)
    gradient = bool(c['gradient']) and cell.params is not None
    eng = build_engine(cell, gradient=gradient)
    progress('sweep', 0, fre.size)
    V = float(c['voltage'])
    def chunk_done(done, total):
        progress('sweep', done, total)
    if gradient:
        Q, dY, info = eng.sweep_charges(fre, V, precision=c['precision'], batch=c['batch'], gradient=True,
                                        on_chunk=chunk_done)
    else:
        Q, info = eng.sweep_charges(fre, V, precision=c['precision'], batch=c['batch'], on_chunk=chunk_done)
        dY = None
    if not np.all(np.isfinite(Q)):
        raise ValueError('non-finite charge in the sweep')
    progress('postprocess', fre.size, fre.size)
    y = -1j * 2 * np.pi * fre * Q / V
    peak = int(np.argmax(np.abs(y)))
    disp, _, _ = eng.solve(2 * np.pi * fre[peak], V)
    n, nd = cell.model.N, cell.ndisp
    u = np.column_stack([disp[i * n:(i + 1) * n] for i in range(nd)]) * SCALE
    phi = disp[nd * n:(nd + 1) * n]
    labels = ['ux', 'uy', 'uz'][:nd]
    fields = dict(position_m=cell.model.pos_m, displacement_m=u, potential_v=phi, frequency_hz=fre[peak],
                  quads=cell.model.quads, mesh_dimension=2, displacement_components=nd, component_labels=labels)
    meta = dict(model_id=model.model_id, nodes=int(cell.model.N), elements=int(cell.model.ne),
                dofs=int(eng.gdof), active_unknowns=int(eng.n_active), pml_reference_frequency_hz=float(fre[-1]),
                backend='MI355X batched multifrontal LU (HIP, libsawfem_hip)', precision=c['precision'],
                refinement=[{k: v for k, v in it.items() if k != 'dq_hist'} for it in info],
                elapsed_seconds=time.time() - started, displacement_components=nd, component_labels=labels,
                field_selection='maximum sampled admittance magnitude')
    write_sweep_results(out, fre, Q, V, mesh=(cell.model.pos_m, cell.model.quads, cell.model.elem_tag),
                        fields=fields, gradient=(prm.NAMES, dY) if dY is not None else None, metadata=meta)
    progress('complete', fre.size, fre.size)
    return Result(out)
