"""The oracle on the rebuilt template meshes reproduces the reference's stored |Y| curves (CPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

import helpers
from sawsim_amd import workloads

DATA = Path(__file__).parent / 'data'


@pytest.mark.parametrize('name', ['sp_double_layer', 'sp_quad_layer'])
def test_oracle_reproduces_stored_curve(name):
    meta = json.loads((DATA / f'{name}.json').read_text())
    d = np.load(DATA / f'{name}.npz')
    idx = np.array(meta['test_indices'])
    f = d['frequency_hz'][idx]
    build = workloads.double_layer if name == 'sp_double_layer' else workloads.quad_layer
    cell = build(points=len(f), band=(f[0], f[-1]))
    eng = helpers.oracle_engine(cell, 2 * np.pi * f[-1])
    y = np.array([abs(eng.solve(2 * np.pi * fz)[2]) for fz in f])
    own = d['sawsim_magnitude'][idx]
    err = np.abs(y - own) / own
    assert err.max() <= meta['regression_rtol']
    assert err.max() <= 1e-7          # the rebuilt mesh is the reference's mesh
