"""fr/fa extraction (host logic, CPU), pinned by the reference's own known-answer tests.

Synthetic curves: a lossless resonator |Y| ~ f (f^2 - fa^2) / (f^2 - fr^2) has its pole and
zero at known frequencies (``tests/test_agent.py:16-52`` of the reference); stored curves: the
fr/fa the reference's validation table lists for its released template curves
(``docs/validation.en.md:10-20``) and the 1 MHz agreement of its ``test_reference_fr_fa_agree``.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from sawsim_amd.metrics import compare, resonances, vertex_fit

DATA = Path(__file__).parent / 'data'


def _lossless(fr, fa, n=201, lo=1.6e9, hi=2.0e9):
    f = np.linspace(lo, hi, n)
    return f, np.abs(f * (f ** 2 - fa ** 2) / (f ** 2 - fr ** 2))


def test_vertex_fit_is_exact_for_a_v():
    f = np.array([0.0, 1.0, 2.0])
    for x0 in (0.7, 1.0, 1.3):
        g = 3.0 + 2.0 * np.abs(f - x0)
        assert vertex_fit(f, g, 1) == pytest.approx(x0, abs=1e-12)


def test_synthetic_pole_and_zero_between_samples():
    fr, fa = 1.7512345e9, 1.8123456e9
    f, y = _lossless(fr, fa)
    r = resonances(f, y)
    assert abs(r['fr_hz'] - fr) < r['uncertainty_hz'] and abs(r['fa_hz'] - fa) < r['uncertainty_hz']
    assert r['k2eff'] == pytest.approx(math.pi ** 2 / 4 * (fa - fr) / fa, rel=1e-2)
    assert not r['warnings']


def test_weak_coupling_on_a_capacitive_background():
    f = np.linspace(2.0e9, 3.0e9, 401)
    fr, fa = 2.47310e9, 2.47685e9
    y = 2j * np.pi * f * 1e-12 * (fa ** 2 - f ** 2) / (fr ** 2 - f ** 2)
    r = resonances(f, y)
    assert abs(r['fr_hz'] - fr) < 2.5e6 and abs(r['fa_hz'] - fa) < 2.5e6
    assert 'peak_at_band_edge' not in r['warnings']


def test_band_edge_and_missing_antiresonance_warnings():
    r = resonances(*_lossless(1.55e9, 1.62e9))
    assert r['fr_hz'] is None and 'peak_at_band_edge' in r['warnings']
    r = resonances(*_lossless(1.9511e9, 2.2e9))
    assert r['fr_hz'] is not None and r['fa_hz'] is None and 'antiresonance_not_found' in r['warnings']


# fr / fa in GHz of the reference curve and of the released solver curve (validation table)
TABLE = {'sp_single_layer': ((1.80791, 1.87557), (1.80785, 1.87557)),
         'sp_double_layer': ((1.82362, 1.90406), (1.82363, 1.90407)),
         'sp_quad_layer': ((1.75859, 1.84382), (1.75859, 1.84382))}


@pytest.mark.parametrize('name', sorted(TABLE))
def test_stored_curves_reproduce_the_validation_table(name):
    z = np.load(DATA / f'{name}.npz')
    ref = resonances(z['frequency_hz'], z['reference_magnitude'])
    sol = resonances(z['frequency_hz'], z['sawsim_magnitude'])
    (rfr, rfa), (sfr, sfa) = TABLE[name]
    # the table rounds to 10 kHz
    assert abs(ref['fr_hz'] * 1e-9 - rfr) <= 6e-6 and abs(ref['fa_hz'] * 1e-9 - rfa) <= 6e-6
    assert abs(sol['fr_hz'] * 1e-9 - sfr) <= 6e-6 and abs(sol['fa_hz'] * 1e-9 - sfa) <= 6e-6
    d = compare(ref, sol)
    assert abs(d['fr_diff_hz']) < 1e6 and abs(d['fa_diff_hz']) < 1e6
