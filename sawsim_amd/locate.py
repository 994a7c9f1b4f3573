"""Coarse sweep -> fr/fa (widening the band when needed) -> zoomed sweep around them.

The reference's ``locate`` (``src/sawsim/agent.py:406-477``) re-runs the whole pipeline for every
band: mesh, assembly, a new engine, a new sparse factorization per point.  Here one GPU engine
serves every band.  Its symbolic analysis and numeric plan are built once; a new band only
re-assembles K, M and D on the GPU (the PML is evaluated at the band's top frequency, so the
operator changes with the band) and re-gathers them into the plan
(:meth:`FrequencySolver.update_operator`).

The band decisions follow the reference, restated:

* no resonance/antiresonance pair in the band, or a peak at the band edge with the band already
  at the limits: sample more densely (2 p + 1 points, at most 401), else stop;
* a peak at a band edge: move that edge out by half the span (within the limits);
* no antiresonance after fr: move the top edge out by half the span (within the limits);
* then zoom to [fr - h, fa + h] with h = max(fa - fr, 2 steps) and ``zoom_points`` points,
  zooming again while fewer than five samples lie between fr and fa (at most three zooms).

Not carried over: the reference's rounding of band edges to 0.1 kHz in its config files, and
its optional Q refinement (``_add_quality``, an agent-level post-processing step).
"""
from __future__ import annotations

from dataclasses import replace
from typing import Callable, Optional

import numpy as np

from . import metrics


def locate_bands(sweep: Callable, band_hz, points: int, limits_hz=None, zoom_points: int = 101,
                 max_rounds: int = 6, max_zooms: int = 3) -> dict:
    """The band logic on any sweep callable: ``sweep(band_hz, points) -> (f_hz, admittance)``.

    Returns ``ok``, ``fr_hz``, ``fa_hz``, ``k2eff``, ``step_hz``, the last curve
    (``frequency_hz``, ``admittance``), ``coarse`` (band and fr/fa of the last full-band sweep)
    and ``history`` (one entry per sweep); ``ok`` False with ``error`` when fr and fa cannot be
    bracketed or the zoom loses them.
    """
    band = (float(band_hz[0]), float(band_hz[1]))
    lo, hi = (float(limits_hz[0]), float(limits_hz[1])) if limits_hz is not None else band
    history = []

    def run(b, n):
        f, y = sweep(b, n)
        r = metrics.resonances(f, y)
        history.append(dict(band_hz=[b[0], b[1]], points=int(n), fr_hz=r['fr_hz'], fa_hz=r['fa_hz'],
                            warnings=list(r['warnings'])))
        return np.asarray(f), np.asarray(y), r

    f, y, r = None, None, None
    for _ in range(max_rounds):
        f, y, r = run(band, points)
        warn = set(r['warnings'])
        span = band[1] - band[0]
        at_limits = band[0] <= lo and band[1] >= hi
        if 'no_resonance_found' in warn or ('peak_at_band_edge' in warn and at_limits):
            if points >= 401:
                break
            points = min(401, 2 * points + 1)       # weak coupling: sample the fr..fa gap
            continue
        if 'peak_at_band_edge' in warn:
            g = np.abs(y) / f
            if int(np.argmax(g)) == 0:
                band = (max(lo, band[0] - span / 2), band[1])
            else:
                band = (band[0], min(hi, band[1] + span / 2))
            continue
        if 'antiresonance_not_found' in warn and band[1] < hi:
            band = (band[0], min(hi, band[1] + span / 2))
            continue
        break
    if r is None or r['fr_hz'] is None or r['fa_hz'] is None:
        return dict(ok=False, error='could not bracket fr and fa within %.6g..%.6g Hz' % (lo, hi), history=history)
    coarse = dict(band_hz=[band[0], band[1]], fr_hz=r['fr_hz'], fa_hz=r['fa_hz'])
    z, zf, zy = None, f, y
    last = r
    for _ in range(max_zooms):
        half = max(last['fa_hz'] - last['fr_hz'], 2 * last['step_hz'])
        zb = (max(lo, last['fr_hz'] - half), min(hi, last['fa_hz'] + half))
        f2, y2, r2 = run(zb, zoom_points)
        if r2['fr_hz'] is None or r2['fa_hz'] is None:
            break
        z, zf, zy, last = r2, f2, y2, r2
        if 'coarse_sampling' not in r2['warnings']:
            break
    if z is None:
        return dict(ok=False, error='the zoomed sweep lost the resonance', coarse=coarse, history=history)
    return dict(ok=True, fr_hz=z['fr_hz'], fa_hz=z['fa_hz'], k2eff=z['k2eff'], step_hz=z['step_hz'],
                uncertainty_hz=z['uncertainty_hz'], warnings=z['warnings'], frequency_hz=zf, admittance=zy,
                coarse=coarse, history=history)


def locate(cell, *, limits_hz=None, zoom_points: int = 101, precision: str = 'mixed', engine=None,
           V: float = 1.0, on_sweep: Optional[Callable] = None) -> dict:
    """:func:`locate_bands` on a workload cell (Q9 stack or Hex27 slab) with one GPU engine.

    ``limits_hz`` bounds every band (default: the cell's band).  ``engine`` may be a prebuilt
    engine of this cell; otherwise one is built once.  ``on_sweep(band_hz, points)`` is called
    before each sweep.  The result also holds the engine and the number of sweeps it ran.
    """
    from . import workloads
    hexcell = isinstance(cell, workloads.HexCell)
    eng = engine if engine is not None else workloads.build_engine(cell)
    sign = 1j if hexcell else -1j                  # Y = +i w Q / V on slabs, -i w Q / V in 2D

    def sweep(band, points):
        if on_sweep is not None:
            on_sweep(band, points)
        c = replace(cell, band_hz=(band[0], band[1]), points=int(points))
        K, M, D = workloads.assemble_hex_cell(c) if hexcell else workloads.assemble_cell(c)
        eng.update_operator(K, M, D, beta=cell.beta_dk)
        f = c.frequencies()
        Q, _ = eng.sweep_charges(f, V, precision=precision)
        return f, sign * 2 * np.pi * f * Q / V

    out = locate_bands(sweep, cell.band_hz, cell.points, limits_hz=limits_hz, zoom_points=zoom_points)
    out['engine'] = eng
    out['sweeps'] = len(out['history'])
    return out
