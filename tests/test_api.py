"""Model/sweep/Result mirror: result-directory layout (CPU) and one GPU sweep."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from sawsim_amd.api import Model, Result, sweep, write_sweep_results


def test_model_fields_are_checked():
    m = Model('sp_double_layer', points=5, precision='fp64')
    assert m.to_dict()['points'] == 5 and 'sp_quad_layer' in Model.templates()
    with pytest.raises(ValueError):
        Model('nope')
    with pytest.raises(ValueError):
        Model('config1', pitch_um=1.0)          # synthetic cells take the sweep fields only
    # the reference's bounds on the sweep fields
    for bad in (dict(points=2), dict(points=1202), dict(points=5.0), dict(points=True), dict(start_ghz=0),
                dict(start_ghz=2.0, stop_ghz=1.5), dict(voltage=-1.0), dict(precision='fp32'), dict(batch=0),
                dict(start_ghz=3.0), dict(voltage=float('inf')), dict(stop_ghz=float('nan')), dict(gradient=1)):
        with pytest.raises(ValueError):
            Model('sp_double_layer', **bad)
    for bad in (dict(start_ghz=1.6), dict(voltage=float('inf')), dict(points=1.0e2)):
        with pytest.raises(ValueError):
            Model('config1', **bad)              # 1.6 GHz is above config 1's default stop (1.5429 GHz)
    Model('sp_double_layer', points=3, start_ghz=1.5, stop_ghz=np.float64(2.7), voltage=2, batch=4)


def test_model_takes_the_reference_simulation_config_fields():
    """Every SimulationConfig field over the template defaults, with the reference's ranges
    (src/sawsim/config.py:16-124, src/sawsim/sp_specs.py:4-80)."""
    m = Model('sp_double_layer', pitch_um=1.2, points=101, mesh_um=0.4, metal_ratio=0.45, electrode_um=0.15,
              substrate_um=0.8, layers=[dict(material_id='si_isotropic', thickness_um=6.0)],
              euler_theta_deg=-38.0, mode_extension=0, beta_dk=1e-13, eta_eps=1e-3)
    c = m.to_dict()
    assert c['pitch_um'] == 1.2 and c['start_ghz'] == 1.5 and c['layers'][0]['euler_phi_deg'] == 0.0
    cell = m.cell()
    assert cell.ndisp == 2 and cell.beta_dk == 1e-13 and cell.points == 101
    assert np.isclose(np.ptp(cell.model.pos_m[:, 0]), 2 * 1.2e-6)
    # a saved reference input (the stored curve's configuration) loads as it is
    cfg = json.loads((Path(__file__).parent / 'data' / 'sp_quad_layer.json').read_text())['config']
    assert Model(**cfg).to_dict()['points'] == cfg['points']
    stack = Model('sp_stack', pitch_um=1.0, substrate_um=0.5, mesh_um=0.1,
                  layers=[dict(material_id='sio2', thickness_um=0.4), dict(material_id='si_isotropic', thickness_um=4)],
                  coatings=[dict(material_id='sio2', thickness_um=0.3)]).cell()
    assert len(stack.materials) == 5 and stack.ndisp == 3
    bad_cases = [
        ('sp_double_layer', dict(pitch_um=1.6)),                     # pitch range [0.8, 1.5]
        ('sp_double_layer', dict(mesh_um=0.1)),                      # mesh range [0.25, 0.8]
        ('sp_double_layer', dict(layers=[])),                        # needs one backing layer
        ('sp_double_layer', dict(coatings=[dict(material_id='sio2', thickness_um=0.5)])),
        ('sp_double_layer', dict(aperture_um=0.2)),                  # 2.5D only
        ('sp_double_layer', dict(beta_dk=2e-10)), ('sp_double_layer', dict(eta_eps=0.2)),
        ('sp_double_layer', dict(mode_extension=2)), ('sp_double_layer', dict(euler_phi_deg=400.0)),
        ('sp_double_layer', dict(metal_ratio=float('nan'))), ('sp_double_layer', dict(substrate_material='linbo3')),
        ('sp_double_layer', dict(layers=[dict(material_id='unobtainium', thickness_um=1.0)])),
        ('sp_double_layer', dict(electrode_material='si_isotropic')),   # not an electrode record
        ('sp_double_layer', dict(foo=1)), ('sp_double_layer', dict(material_snapshots={'x': {}})),
        ('sp_stack', dict(coatings=[dict(material_id='sio2', thickness_um=0.1)])),   # must embed the electrode
        ('sp_stack', dict(layers=[dict(material_id='sio2', thickness_um=0.1)] * 7)),
        ('sp_2p5d_double_layer', dict(beta_dk=1e-13)),               # loss on 2D templates only
        ('sp_2p5d_double_layer', dict(aperture_um=None)), ('sp_2p5d_quad_layer', dict(mode_extension=0)),
        ('sp_single_layer', {}), ('sp_tcsaw', {}),                   # unstructured Gmsh meshes: refused
    ]
    for mid, bad in bad_cases:
        with pytest.raises(ValueError):
            Model(mid, **bad)


def test_template_cells_equal_the_stored_curve_cells():
    """The default fields of sp_double_layer / sp_quad_layer / sp_2p5d_* build exactly the cells
    the stored-curve tests run on (mesh, material index, tensors)."""
    from sawsim_amd import workloads
    pairs = [('sp_double_layer', workloads.double_layer()), ('sp_quad_layer', workloads.quad_layer())]
    for mid, ref in pairs:
        got = Model(mid, points=5).cell()
        assert np.array_equal(got.model.pos_m, ref.model.pos_m) and np.array_equal(got.model.quads, ref.model.quads)
        assert np.array_equal(got.mat_idx, ref.mat_idx) and got.ndisp == ref.ndisp
        for a, b in zip(got.materials, ref.materials):
            assert a.rho == b.rho and np.array_equal(a.C, b.C) and np.array_equal(a.e, b.e)
            assert np.array_equal(a.eps, b.eps)
    for name in ('single', 'quad'):
        ref = workloads.hex_template(name)
        got = Model(f'sp_2p5d_{name}_layer', points=5).cell()
        assert np.array_equal(got.mesh.pos_m, ref.mesh.pos_m) and np.array_equal(got.mesh.hexas, ref.mesh.hexas)
        for a, b in zip(got.materials, ref.materials):
            assert np.array_equal(a.C, b.C) and np.array_equal(a.eps, b.eps)
        assert got.band_hz == ref.band_hz


def test_result_directory_round_trip(tmp_path):
    f = np.linspace(1.6e9, 2.0e9, 201)
    fr, fa = 1.7512345e9, 1.8123456e9           # between samples
    y_true = 2j * np.pi * f * 1e-12 * (fa ** 2 - f ** 2) / (fr ** 2 - f ** 2)
    q = y_true / (-1j * 2 * np.pi * f)            # Y = -i w Q / V with V = 1
    pos = np.zeros((9, 2))
    write_sweep_results(tmp_path, f, q, 1.0, mesh=(pos, np.arange(9)[None], np.array([16])),
                        gradient=(['a', 'b'], np.ones((f.size, 2))), metadata=dict(model_id='synthetic'))
    r = Result(tmp_path)
    assert np.allclose(r.frequency_ghz, f * 1e-9) and np.allclose(r.admittance, y_true)
    assert r.peak_frequency_ghz == pytest.approx(f[np.argmax(np.abs(y_true))] * 1e-9)
    names = {a['name'] for a in r.manifest['artifacts']}
    assert {'curve.json', 'admittance.npz', 'admittance.csv', 'mesh.npz', 'gradient.npz', 'metadata.json',
            'manifest.json'} <= names
    z = np.load(tmp_path / 'admittance.npz')
    assert np.allclose(z['admittance'], y_true) and np.allclose(z['charge'], q)
    csv = np.loadtxt(tmp_path / 'admittance.csv', delimiter=',', skiprows=1)
    assert np.allclose(csv[:, 3], np.abs(y_true))
    res = r.resonances()
    assert abs(res['fr_hz'] - fr) < res['uncertainty_hz'] and abs(res['fa_hz'] - fa) < res['uncertainty_hz']
    assert r.save(tmp_path / 'copy').joinpath('curve.json').exists()


@pytest.mark.gpu
def test_sweep_reproduces_the_stored_double_layer_curve(tmp_path):
    print('[test_sweep_reproduces_the_stored_double_layer_curve]', flush=True)
    sys.stdout.flush()
    from pathlib import Path
    r = sweep(Model('sp_double_layer', points=5), tmp_path / 'run')
    z = np.load(Path(__file__).parent / 'data' / 'sp_double_layer.npz')
    ref = z['sawsim_magnitude'][::200]
    assert np.allclose(z['frequency_hz'][::200] * 1e-9, r.frequency_ghz)
    assert np.max(np.abs(r.magnitude - ref) / ref) <= 2e-3
    meta = json.loads((tmp_path / 'run' / 'metadata.json').read_text())
    assert meta['model_id'] == 'sp_double_layer' and meta['frequency_count'] == 5
    f = np.load(tmp_path / 'run' / 'fields.npz')
    assert f['displacement_m'].shape[1] == 3 and np.all(np.isfinite(f['potential_v']))
    with pytest.raises(FileExistsError):
        sweep(Model('sp_double_layer', points=5), tmp_path / 'run')


def test_mesh_only_writes_the_mesh_and_stops(tmp_path):
    out = tmp_path / 'mesh'
    r = sweep(Model('sp_double_layer', points=5), out, mesh_only=True)
    meta = json.loads((out / 'metadata.json').read_text())
    assert meta['task_kind'] == 'mesh' and meta['model_id'] == 'sp_double_layer' and meta['nodes'] > 0
    m = np.load(out / 'mesh.npz')
    assert m['quads'].shape[1] == 9 and m['position_m'].shape == (meta['nodes'], 2)
    assert m['material_tags'].shape[0] == meta['elements']
    names = {a['name'] for a in r.manifest['artifacts']}
    assert {'input.json', 'mesh.npz', 'metadata.json', 'manifest.json'} <= names
    assert not (out / 'curve.json').exists() and r.frequency_ghz.size == 0


def test_cancellation_raises_before_any_solve(tmp_path):
    events = []
    with pytest.raises(InterruptedError):
        sweep(Model('sp_double_layer', points=5), tmp_path / 'a', on_progress=events.append,
              should_cancel=lambda: True)
    assert events == [] and not (tmp_path / 'a' / 'curve.json').exists()
    (tmp_path / 'b').mkdir()
    (tmp_path / 'b' / 'cancel').write_text('')
    with pytest.raises(InterruptedError):
        sweep(Model('sp_double_layer', points=5), tmp_path / 'b', mesh_only=True)


@pytest.mark.gpu
def test_cancellation_between_frequency_chunks(tmp_path):
    print('[test_cancellation_between_frequency_chunks]', flush=True)
    sys.stdout.flush()
    events = []

    def cancel():
        return any(e['stage'] == 'sweep' and e['completed'] > 0 for e in events)
    with pytest.raises(InterruptedError):
        sweep(Model('sp_double_layer', points=9, batch=3), tmp_path / 'run', on_progress=events.append,
              should_cancel=cancel)
    done = [e['completed'] for e in events if e['stage'] == 'sweep']
    assert done[-1] == 3 and not (tmp_path / 'run' / 'curve.json').exists()
