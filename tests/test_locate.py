"""Band logic of ``locate`` (host side, CPU) on an analytic resonator admittance.

Y(f) = i w C0 (fa^2 - f^2 + i g) / (fr^2 - f^2 + i g): a pole at fr (|Y| maximum) and a zero at
fa (|Y| minimum) on the static-capacitance slope, the curve shape ``metrics.resonances`` is built
for; the true fr/fa are known, so the zoomed estimates are checked against them.
"""
import numpy as np

from sawsim_amd.locate import locate_bands

FR, FA = 1.80e9, 1.86e9


def resonator(g=1e15):
    calls = []

    def sweep(band, points):
        calls.append((band, points))
        f = np.linspace(band[0], band[1], points)
        w = 2 * np.pi * f
        return f, 1j * w * 1e-12 * (FA ** 2 - f ** 2 + 1j * g) / (FR ** 2 - f ** 2 + 1j * g)
    return sweep, calls


def test_zoom_refines_fr_and_fa_inside_the_band():
    sweep, calls = resonator()
    out = locate_bands(sweep, (1.6e9, 2.0e9), 41, zoom_points=101)
    assert out['ok'] and len(calls) >= 2
    zb, zp = calls[-1]
    assert zp == 101 and zb[0] < FR < FA < zb[1] and zb[1] - zb[0] < 0.4e9
    # the zoom step is (zb[1] - zb[0]) / 100; the vertex fit lands within a step of the truth
    step = (zb[1] - zb[0]) / 100
    assert abs(out['fr_hz'] - FR) < step and abs(out['fa_hz'] - FA) < step
    assert out['coarse']['band_hz'] == [1.6e9, 2.0e9]


def test_band_below_the_resonance_widens_upwards():
    sweep, calls = resonator()
    out = locate_bands(sweep, (1.2e9, 1.6e9), 41, limits_hz=(1.0e9, 3.0e9))
    assert out['ok']
    bands = [b for b, _ in calls]
    assert bands[1][0] == 1.2e9 and bands[1][1] > 1.6e9          # top edge moved out by half a span
    assert out['coarse']['band_hz'][1] > FA
    assert abs(out['fr_hz'] - FR) < 5e6 and abs(out['fa_hz'] - FA) < 5e6


def test_no_resonance_within_the_limits_is_reported():
    sweep, calls = resonator()
    out = locate_bands(sweep, (0.5e9, 0.9e9), 41, limits_hz=(0.5e9, 0.9e9))
    assert not out['ok'] and 'error' in out
    assert max(p for _, p in calls) <= 401
