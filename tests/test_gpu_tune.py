"""GPU: the kernel-selection tuner picks among its candidates, leaves the engine set to its
choice, and the sweep after tuning gives the same Y11 as before (to the mixed path's accuracy)."""
import sys

import numpy as np
import pytest

from sawsim_amd import tune, workloads

pytestmark = pytest.mark.gpu


def test_autotune_choice_and_results():
    print('[test_autotune_choice_and_results]', flush=True)
    sys.stdout.flush()
    cell = workloads.config3(mesh_um=0.09, points=12)
    eng = workloads.build_engine(cell, gradient=True)
    f = cell.frequencies()
    Q0, dY0, _ = eng.sweep_charges(f, precision='mixed', gradient=True)
    res = tune.autotune(eng, f, 'mixed', gradient=True, batch=6, reps=1)
    print(res, flush=True)
    default = tune._defaults()
    for knob, v in res['choice'].items():
        assert set(res['ms'][knob]) == set(tune.CANDIDATES[knob]) | {default[knob]}
        assert v == default[knob] or res['ms'][knob][v] < 0.98 * res['ms'][knob][default[knob]]
        assert all(ms > 0 for ms in res['ms'][knob].values())
    Q1, dY1, _ = eng.sweep_charges(f, precision='mixed', gradient=True)
    Q64, _ = eng.sweep_charges(f, precision='fp64')
    assert np.max(np.abs(Q1 - Q64) / np.abs(Q64)) <= 1e-6 and np.max(np.abs(Q0 - Q64) / np.abs(Q64)) <= 1e-6
    scale = np.abs(cell.params)
    scale[12:15] = 1.0
    assert np.max(np.abs((dY1 - dY0) * scale)) <= 1e-5 * np.max(np.abs(dY0 * scale))
