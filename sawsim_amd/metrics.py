"""Resonance fr and antiresonance fa of a sampled admittance curve (host side, NumPy).

The fr/fa check of a sweep uses the reference's published extraction method
(``src/sawsim/metrics.py:18-31, 98-171``), restated here:

* work on g = |Y| / f, which removes the static-capacitance slope |Y| ~ 2 pi f C0;
* candidate pairs: every local maximum p of g, with the lowest g between p and the next
  maximum (or the end of the band) as its antiresonance a; pairs whose minimum lies more than
  15 % of f_p above f_p are broad humps and dropped (unless the minimum is the last sample);
* main pair: the largest relative dip (g_p - g_a) / g_p, capped at 1/2 so that every
  well-resolved pair competes on its peak height g_p instead;
* fr and fa are refined between samples by a V fit (two lines of opposite slope through the
  extreme sample and its neighbours; exact for an isolated pole or zero on a flat background):
  fr on 1/g around p, fa on g around a;
* k2eff = pi^2/4 (fa - fr) / fa; the uncertainty is half the frequency step.
"""
from __future__ import annotations

import math

import numpy as np

K2_FORMULA = 'k2eff = pi^2/4 * (fa - fr) / fa'
BROAD_HUMP = 0.15        # fa - fr above this fraction of fr: not a resonator pole/zero pair
DIP_CAP = 0.5            # relative dips above this count as equally well resolved


def vertex_fit(f, g, i):
    """Abscissa of the vertex of the V through (f[i-1..i+1], g[i-1..i+1]), g[i] the minimum.

    The V's slope magnitude is that of the steeper side; the vertex stays within half a step
    of f[i].  Band-edge samples and flat triples return f[i].
    """
    if i <= 0 or i >= len(g) - 1:
        return float(f[i])
    half_span = 0.5 * (f[i + 1] - f[i - 1])
    slope = max(g[i - 1] - g[i], g[i + 1] - g[i]) / half_span
    if not slope > 0:
        return float(f[i])
    shift = (g[i - 1] - g[i + 1]) / (2.0 * slope)
    return float(f[i] + min(max(shift, -0.5 * half_span), 0.5 * half_span))


def local_maxima(y):
    """Indices i of interior samples with y[i-1] < y[i] >= y[i+1]."""
    return np.flatnonzero((y[1:-1] > y[:-2]) & (y[1:-1] >= y[2:])) + 1


def resonances(frequency_hz, admittance):
    """fr/fa/k2eff of a curve; ``admittance`` is complex Y or |Y|, frequencies in Hz.

    Returns a dict with fr_hz, fa_hz (None when not found), k2eff, the frequency step, the
    uncertainty (half a step) and a list of warnings.
    """
    f = np.asarray(frequency_hz, dtype=float)
    mag = np.abs(np.asarray(admittance))
    n = f.size
    out = dict(fr_hz=None, fa_hz=None, k2eff=None, k2_formula=K2_FORMULA, step_hz=None, uncertainty_hz=None,
               warnings=[])
    if n < 5 or not np.all(np.isfinite(mag)):
        out['warnings'].append('curve too short or not finite')
        return out
    step = float(np.max(np.diff(f)))
    out['step_hz'] = step
    out['uncertainty_hz'] = 0.5 * step
    g = mag / f
    peaks = local_maxima(g)
    pairs = []                                       # (relative dip, peak index, dip index)
    for k, p in enumerate(peaks):
        stop = int(peaks[k + 1]) if k + 1 < len(peaks) else n - 1
        a = int(p) + int(np.argmin(g[p:stop + 1]))
        if a == p:
            continue
        if a != n - 1 and f[a] - f[p] > BROAD_HUMP * f[p]:
            continue
        pairs.append(((g[p] - g[a]) / g[p], int(p), a))
    if not pairs:
        edge = int(np.argmax(g)) in (0, n - 1)
        out['warnings'].append('peak_at_band_edge' if edge else 'no_resonance_found')
        return out
    _, r, a = max(pairs, key=lambda t: (min(t[0], DIP_CAP), g[t[1]]))
    inv = 1.0 / np.maximum(g, np.finfo(float).tiny)
    fr = vertex_fit(f, inv, r)
    out['fr_hz'] = fr
    if a == n - 1:
        out['warnings'].append('antiresonance_not_found')
        return out
    fa = vertex_fit(f, g, a)
    out['fa_hz'] = fa
    out['k2eff'] = math.pi ** 2 / 4 * (fa - fr) / fa
    if a - r < 6:
        out['warnings'].append('coarse_sampling')
    return out


def compare(reference, candidate):
    """fr/fa differences (Hz) of two ``resonances`` results (candidate - reference)."""
    d = {}
    for key in ('fr_hz', 'fa_hz'):
        a, b = reference.get(key), candidate.get(key)
        d[key.replace('_hz', '_diff_hz')] = None if a is None or b is None else b - a
    return d
