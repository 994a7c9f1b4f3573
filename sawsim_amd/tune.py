"""Kernel-selection tuner of the batched multifrontal solver.

The engine has one algorithmic knob, ``share`` (consecutive frequencies per complex64
factorization, :meth:`sweep.FrequencySolver.solve_batch`), whose best value depends on how far
apart the sweep's frequencies are relative to the cell's resonances: :func:`tune_share` times
each candidate on a block of the sweep's own frequencies.  The plan has three kernel-selection
knobs whose best setting depends on the cell's front sizes and on the chunk size:

* ``panel_lu`` (factorization): 32-wide pivot panels of size buckets up to this front order are
  factored by one kernel per panel instead of the diagonal-block kernel and two MFMA GEMMs;
* ``solve_full`` (solves): fronts up to this order are solved by one workgroup each;
* ``tri_lds`` (solves): one-workgroup buckets whose factor block fits this many bytes are solved
  from LDS by persistent workgroups.

:func:`autotune` times the candidates on one chunk of the actual sweep, with HIP events on the
launch stream, one knob at a time (coordinate descent: the factorization knob on the
factorization, the solve knobs on a forward + backward pass with the chunk's right-hand sides),
keeps the fastest setting of each, and records timings and choice in ``eng.tuning``.  The
results of the solver do not depend on the knobs beyond rounding (the GPU tests compare the
paths); only the time does.
"""
from __future__ import annotations

import numpy as np

from . import _lib

CANDIDATES = {
    'panel_lu': (0, 128, 256, 512),
    'solve_full': (128, 192, 256, 384),
    'tri_lds': (0, 16384, 32768),
}
_SETTERS = {'panel_lu': 'set_panel_lu', 'solve_full': 'set_solve_full', 'tri_lds': 'set_tri_lds'}


def _defaults():
    from . import sweep
    return dict(panel_lu=sweep.PANEL_LU_M, solve_full=sweep.SOLVE_FULL_M, tri_lds=sweep.TRI_LDS_BYTES)


def autotune(eng, freqs, precision='mixed', gradient=True, batch=None, reps=3, candidates=None, margin=0.02):
    """Pick the fastest setting of each knob for chunks of the sweep over ``freqs`` (Hz).

    ``batch``: frequencies of the timed chunk (default: what one chunk of the sweep holds).  A
    candidate replaces the library default only when it is faster by more than ``margin`` (run to
    run noise of these timings is about 1 %).  Returns ``dict(choice=..., ms=...)`` (also kept in
    ``eng.tuning``); the engine is left set to the choice.
    """
    from .piezo_fem import stream_ptr
    torch = eng.torch
    cand = dict(CANDIDATES, **(candidates or {}))
    freqs = np.asarray(freqs, dtype=float).ravel()
    B = int(min(freqs.size, batch or eng.auto_batch(precision, gradient)))
    B = -(-freqs.size // -(-freqs.size // B))           # the equal chunks sweep_charges uses
    prec = 0 if precision == 'mixed' else 1
    nb = 64 if prec == 0 else 32
    R = 2 if gradient else 1
    n = eng.n_active
    om = torch.as_tensor(2 * np.pi * freqs[:B], device=eng.dev)
    arena = eng._arena(eng.mf.arena_elems * B, torch.complex64 if prec == 0 else torch.complex128)
    uarena = torch.empty(max(1, eng.mf.uarena_elems * B * R), dtype=torch.complex128, device=eng.dev)
    gen = torch.Generator(device=eng.dev).manual_seed(0)
    x0 = torch.randn((B, R, n), dtype=torch.complex128, device=eng.dev, generator=gen)
    x = torch.empty_like(x0)

    def timed(fn):
        best = None
        for _ in range(reps + 1):                     # the first call warms up (work lists, code)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None or ms < best else best
        return best

    def factor():
        _lib.call('sawfem_mf_factor', eng._plan, stream_ptr(), prec, B, om.data_ptr(), eng.beta, eng.Kaa.data_ptr(),
                  eng.Maa.data_ptr(), eng._dptr(eng.Daa), arena.data_ptr(), nb)

    def solve():
        x.copy_(x0)
        _lib.call('sawfem_mf_solve', eng._plan, stream_ptr(), prec, B, R, arena.data_ptr(), uarena.data_ptr(),
                  x.data_ptr(), 0, nb, None, 0)

    choice, ms = {}, {}
    default = _defaults()
    for knob, fn in (('panel_lu', factor), ('solve_full', solve), ('tri_lds', solve)):
        setter = getattr(eng, _SETTERS[knob])
        ms[knob] = {}
        values = list(dict.fromkeys((default[knob],) + tuple(cand[knob])))
        for v in values:
            setter(v)
            ms[knob][v] = timed(fn)
        best = min(ms[knob], key=ms[knob].get)
        choice[knob] = best if ms[knob][best] < (1 - margin) * ms[knob][default[knob]] else default[knob]
        setter(choice[knob])
        if knob == 'panel_lu':
            factor()                                 # the arena the solve knobs are timed on
    torch.cuda.synchronize()
    eng.pivot_stats(reset=True)                      # the tuning factorizations are not the sweep's
    eng.tuning = dict(choice=choice, ms=ms, batch=B)
    return eng.tuning


SHARE_CANDIDATES = (1, 2, 4, 8, 16)


def tune_share(eng, freqs, precision='mixed', gradient=True, candidates=SHARE_CANDIDATES, margin=0.02, reps=2):
    """Pick the factorization sharing of the sweep over ``freqs`` (Hz).  Each candidate solves the
    chunk it would run in the sweep: as many of the sweep's frequencies as fit the device memory
    at that sharing (more with more sharing: the arena holds one factorization per ``share``
    frequencies), a contiguous block from the middle of the band, refinement and fallbacks
    included, timed with the device synchronized (the fastest of ``reps`` runs).  The choice is
    the smallest sharing whose time per frequency is within ``margin`` of the fastest candidate's
    (less sharing takes fewer refinement steps, so near-ties go to it and the choice does not
    flip between runs; in particular share 1 stays unless another is faster by more than
    ``margin``).  Returns ``dict(choice=..., ms=..., batch=...)``; the engine is left set to the
    choice."""
    import time
    torch = eng.torch
    freqs = np.asarray(freqs, dtype=float).ravel()
    P = freqs.size
    candidates = tuple(dict.fromkeys((1,) + tuple(candidates)))
    ws = getattr(eng, '_ws', None)
    keep = 0 if ws is None else ws.numel()
    del ws                       # a reference here would keep the workspace alive when it grows
    from .sweep import hbm_available
    budget = hbm_available(keep)
    batch = {c: int(min(P, eng.auto_batch(precision, gradient, budget=budget, share=c))) for c in candidates}
    # the largest chunk arena up front, so that no candidate's timing includes growing it
    eng._arena(max(eng.chunk_arena_bytes(b, precision, c) for c, b in batch.items()) // 8, torch.complex64)
    ms, per = {}, {}
    for c in candidates:
        B = batch[c]
        lo = max(0, min(P - B, P // 2 - B // 2))
        om = 2 * np.pi * freqs[lo:lo + B]
        # warm-up with this chunk's vector sizes (one factorization for all of it, one step), so
        # that the timed call finds the allocator's blocks and the kernels' code in place
        eng.solve_batch(om, precision=precision, gradient=gradient, share=B, max_ir=1, fallback=False)
        torch.cuda.synchronize()
        for _ in range(reps):                       # the fastest of reps runs
            t0 = time.perf_counter()
            eng.solve_batch(om, precision=precision, gradient=gradient, share=c)
            torch.cuda.synchronize()
            t = 1e3 * (time.perf_counter() - t0)
            ms[c] = min(ms.get(c, t), t)
        per[c] = ms[c] / B
    fastest = min(per.values())
    choice = min(c for c in per if per[c] <= (1 + margin) * fastest)
    eng.set_share(choice)
    eng.pivot_stats(reset=True)
    return dict(choice=choice, ms=ms, batch=batch, ms_per_frequency=per)
