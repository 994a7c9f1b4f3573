"""``locate`` on a GPU engine: coarse sweep, zoom around fr/fa, one engine for every band."""
import sys
from dataclasses import replace

import numpy as np
import pytest

from sawsim_amd import metrics, workloads
from sawsim_amd.locate import locate

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_locate_zooms_on_one_engine_and_matches_a_dense_sweep():
    _say('[test_locate_zooms_on_one_engine_and_matches_a_dense_sweep]')
    cell = workloads.config3(mesh_um=0.12, points=41)
    bands = []
    out = locate(cell, on_sweep=lambda b, p: bands.append((b, p)))
    assert out['ok'], out.get('error')
    assert out['sweeps'] >= 2 and len(bands) == out['sweeps']
    (zb, zp) = bands[-1]
    assert zp == 101
    assert cell.band_hz[0] <= zb[0] < out['fr_hz'] < out['fa_hz'] < zb[1] <= cell.band_hz[1]
    # a dense sweep of the zoom band on an engine built for that band alone puts fr/fa within
    # half a zoom step of the located values
    z = replace(cell, band_hz=(zb[0], zb[1]), points=401)
    eng = workloads.build_engine(z)
    f = z.frequencies()
    Q, _ = eng.sweep_charges(f, precision='mixed')
    r = metrics.resonances(f, -1j * 2 * np.pi * f * Q)
    step = (zb[1] - zb[0]) / (zp - 1)
    _say('located', out['fr_hz'], out['fa_hz'], 'dense', r['fr_hz'], r['fa_hz'], 'zoom step', step)
    assert abs(r['fr_hz'] - out['fr_hz']) <= 0.5 * step
    assert abs(r['fa_hz'] - out['fa_hz']) <= 0.5 * step
