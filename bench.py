#!/usr/bin/env python
"""Frequency-sweep throughput of the MI355X piezo-FEM backend.

One step = one full sweep of the workload: the frequency points factored in groups of
neighbouring frequencies (one complex FP32 factorization per group on MFMA, replayed as a HIP
graph, in chunks that fit HBM), every point solved and refined against its own FP64 operator,
reduced to Y11 and dY11/dp.  Assembly, Bloch reduction and the symbolic analysis are done once
per cell before the timed region (inputs resident in HBM).

Printed JSON line: metric/value (whole-job frequency points per second), the plain FP64
GPU path (one frequency per launch) measured on a sample of points, the Y11 error of the
mixed path against the FP64 path, a roofline entry for the dominant kernel (the grouped
block-solve kernel against HBM, timed live with HIP events), one for the MFMA update GEMM
(`mfma_roofline`) and a CPU baseline (the SuperLU oracle on a bounded sample of the same
workload: one frequency per worker process).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_FP32_MFMA = 157.3   # TFLOP/s dense, MI355X
PEAK_FP64_MFMA = 78.6
# HBM bytes per launch of the grouped solve kernel k_grp_lds (whole fronts: the four instances
# over their dispatches) and of the Schur-complement GEMM k_gemm<float, 1>, each from two rocprofv3
# --pmc runs (FETCH_SIZE, then WRITE_SIZE, calibrated on a 1 GiB device copy), with the
# algorithmic bytes per launch the benchmark reported in the same command and the factorization
# sharing it ran at.  The profiles/ summaries hold the commands and the per-kernel counts.  The
# counters cannot be read inside a timed run (a --pmc run serializes the kernels), so the figures
# are copied from those summaries.  Launch sizes follow the sharing (a launch covers every
# factorization or vector block of a chunk): at another sharing the entry is the measured ratio
# of HBM to algorithmic bytes times the run's own algorithmic bytes per launch.
COUNTER_TRAFFIC = {
    ('grp', 'config2'): dict(share=16, hbm=8.113e8, alg=6.786e8,
                             source='profiles/config2_grouped_solve_traffic.md'),
    ('schur', 'config2'): dict(share=16, hbm=3.562e9, alg=1.848e9, source='profiles/config2_schur_traffic.md'),
    ('schur', 'config4'): dict(share=1, hbm=1.620e10, alg=8.48e9, source='profiles/config4_schur_traffic.md'),
}


def counter_traffic(kernel, workload, share, alg_per_launch):
    """(HBM bytes per launch, note) of COUNTER_TRAFFIC for this run, or (None, None)."""
    e = COUNTER_TRAFFIC.get((kernel, workload))
    if e is None:
        return None, None
    note = 'HBM bytes per launch, rocprofv3 FETCH_SIZE+WRITE_SIZE scaled by a 1 GiB copy'
    if share == e['share']:
        return e['hbm'], f"{note}, from {e['source']} (share {share}, as this run)"
    ratio = e['hbm'] / e['alg']
    return ratio * alg_per_launch, (f"{note}: {e['source']} measured {ratio:.2f}x the algorithmic bytes at share "
                                    f"{e['share']}; times this run's algorithmic bytes per launch (share {share})")


def grp_roofline(grp, workload, share):
    """Roofline entry of the grouped block-solve kernel k_grp_lds (every block size and kind on the
    whole fronts of a grouped solve), the sweep's largest kernel by time: algorithmic bytes (factor
    entries read once per block of vectors, vectors gathered and written; counted by the library per
    launch) over its HIP-event time in the timed region, against HBM."""
    if not grp['launches']:
        return None
    gbs = grp['bytes'] / (grp['ms'] * 1e-3) / 1e9
    traffic, note = counter_traffic('grp', workload, share, grp['bytes'] / grp['launches'])
    return dict(bound='hbm', achieved=gbs, peak=8000.0, unit='GB/s', frac=gbs / 8000.0,
                traffic=traffic, traffic_note=note,
                kernel='k_grp_lds<kind, 8|16, false, true> (grouped whole-front solves on row-major vectors, '
                       'complex64 factors, FP64 vectors)',
                launches=grp['launches'], kernel_ms=grp['ms'], algorithmic_bytes=grp['bytes'],
                algorithmic_bytes_per_launch=grp['bytes'] / grp['launches'],
                avg_launch_ms=grp['ms'] / grp['launches'],
                note='latency-bound: each front is a chain of dependent loads (gather, diagonal blocks, update '
                     'rows, scatter)')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_engine(cell, gradient=True, **kw):
    from sawsim_amd.workloads import build_engine as _build
    return _build(cell, gradient=gradient, **kw)


KINDS = ['init', 'entries', 'diag', 'trsm', 'gemm_l', 'gemm_s', 'fwd', 'bwd', 'small']


def profile_read(eng):
    """Per-kind event totals of the recorded launches, and the same per tree level."""
    import ctypes as C
    from sawsim_amd import _lib
    ms = (C.c_double * 9)()
    cnt = (C.c_int64 * 9)()
    _lib.call('sawfem_mf_profile_read', eng._plan, C.cast(ms, C.c_void_p), C.cast(cnt, C.c_void_p))
    nlev = eng.mf.lvl_ptr.size - 1
    lv = (C.c_double * (9 * nlev))()
    _lib.call_value('sawfem_mf_profile_levels', eng._plan, C.cast(lv, C.c_void_p), nlev)
    per_level = [{n: round(float(lv[l * 9 + i]), 2) for i, n in enumerate(KINDS) if lv[l * 9 + i] > 0}
                 for l in range(nlev)]
    return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(KINDS)}, per_level


def _oracle_engine(cell):
    from sawsim_amd.workloads import HexCell
    if isinstance(cell, HexCell):
        from oracle import fem3d
        mesh = cell.mesh
        M, K = fem3d.assemble(mesh.hexas, mesh.pos_m, cell.material_list(), cell.pml_list(),
                              2 * np.pi * cell.band_hz[1])
        m, el = fem3d.boundary_dofs(mesh.sets, mesh.pos_m, 3 * mesh.N)
        eng = fem3d.Solver3D(K, M, m, el)
        eng.n_active = eng.active.size
        return eng
    sys.path.insert(0, str(ROOT / 'tests'))
    import helpers
    return helpers.oracle_engine(cell, 2 * np.pi * cell.band_hz[1])


def cpu_baseline(cell, workers=8):
    """The oracle (SciPy SuperLU, one complex sparse LU + solve per frequency, the reference's
    per-frequency direct solve) on a bounded sample of the workload: one frequency per worker
    of a fork pool, ``workers`` frequencies spread over the band."""
    from oracle import sweep as osweep
    t0 = time.time()
    eng = _oracle_engine(cell)
    setup = time.time() - t0
    f = cell.frequencies()
    k = max(1, min(workers, os.cpu_count() or 1, f.size))
    sample = f[np.linspace(0, len(f) - 1, k).astype(int)]
    from threadpoolctl import threadpool_limits
    t0 = time.time()
    with threadpool_limits(limits=1):          # one BLAS thread per worker: `cores` = workers
        osweep.sweep_charges(eng, sample, workers=k, complex_form=True)
    dt = time.time() - t0
    return dict(value=k / dt, unit='freq pts/s (Y11 only)', cores=k, kind='port',
                sample=f'{k} of {len(f)} frequencies, one per process of a fork pool of {k} '
                       f'(single-threaded SciPy SuperLU on the complex system, n={eng.n_active}); '
                       f'{dt:.1f}s wall, host setup {setup:.1f}s excluded; os.cpu_count()={os.cpu_count()}')


def plain_fp64(eng, cell, freqs, pp):
    """The plain FP64 GPU path, one frequency per factorization, on pp points of the sweep:
    Y11 only, Y11 with its FP64 adjoint, and Y11 with central differences for the P parameters
    (2P+1 operator rebuilds and solves, measured on one point).  Points per second each."""
    import torch
    from sawsim_amd import workloads
    if isinstance(cell, workloads.HexCell):
        assemble, for_params = workloads.assemble_hex_cell, workloads.hex_materials_for_params
    else:
        assemble, for_params = workloads.assemble_cell, workloads.materials_for_params
    pidx = np.linspace(0, freqs.size - 1, pp).astype(int)
    eng.sweep_charges(freqs[pidx[:1]], precision='fp64', batch=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.sweep_charges(freqs[pidx], precision='fp64', batch=1)
    torch.cuda.synchronize()
    plain = pp / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    eng.sweep_charges(freqs[pidx], precision='fp64', batch=1, gradient=True)
    torch.cuda.synchronize()
    plain_adj = pp / (time.perf_counter() - t0)
    fz = freqs[pidx[:1]]
    p0 = cell.params.copy()
    t0 = time.perf_counter()
    eng.sweep_charges(fz, precision='fp64', batch=1)
    for k in range(eng.n_params):
        h = 1e-6 * abs(p0[k]) if p0[k] != 0 else 1e-6 * 180 / np.pi
        for sgn in (1, -1):
            p = p0.copy()
            p[k] += sgn * h
            K, M, D = assemble(cell, for_params(cell, p))
            eng.update_operator(K, M, D, beta=p[15])
            eng.sweep_charges(fz, precision='fp64', batch=1)
    torch.cuda.synchronize()
    plain_fd = 1.0 / (time.perf_counter() - t0)
    K, M, D = assemble(cell)
    eng.update_operator(K, M, D, beta=p0[15])
    return plain, plain_adj, plain_fd


def assembly_roofline(eng, cell, reps=3):
    """Achieved HBM GB/s of the assembly (K1 element kernel, K2 node-pair gather) and the Bloch
    gathers (K3), re-run ``reps`` times on the built engine with HIP events on the launch stream:
    algorithmic bytes (element matrices written once, read once by the gather, CSR values written,
    index lists read) over the kernel time, against the 8 TB/s HBM peak."""
    import torch
    from sawsim_amd import piezo_fem, workloads
    hexc = isinstance(cell, workloads.HexCell)
    assemble = workloads.assemble_hex_cell if hexc else workloads.assemble_cell
    piezo_fem.KERNEL_LOG = []
    try:
        for _ in range(reps):
            K, M, D = assemble(cell)
            eng.update_operator(K, M, D, beta=eng.beta)
            del K, M, D
        torch.cuda.synchronize()
        out = {}
        for name, a, b, nbytes in piezo_fem.KERNEL_LOG:
            o = out.setdefault(name, dict(ms=0.0, bytes=0, launches=0))
            o['ms'] += a.elapsed_time(b)
            o['bytes'] += nbytes
            o['launches'] += 1
    finally:
        piezo_fem.KERNEL_LOG = None
    for o in out.values():
        o['achieved_gbs'] = o['bytes'] / (o['ms'] * 1e-3) / 1e9
        o['frac'] = o['achieved_gbs'] / 8000.0
        o['ms_per_launch'] = o['ms'] / o['launches']
        o['bytes_per_launch'] = o['bytes'] / o['launches']
    return out


def _mem():
    import torch
    free, total = torch.cuda.mem_get_info()
    return (f'free {free / 1e9:.1f} of {total / 1e9:.1f} GB, torch allocated {torch.cuda.memory_allocated() / 1e9:.1f} '
            f'reserved {torch.cuda.memory_reserved() / 1e9:.1f} GB')


def design_set(args):
    """Config 5: the design set of 64 seeded stack cells (2k-20k Q9 elements, 128 points each,
    Y11 + dY11/dp for 17 parameters), swept back to back on one shared HBM workspace after every
    cell's assembly and symbolic analysis.  Per-cell latency percentiles beside the aggregate."""
    import torch
    from sawsim_amd import _lib, workloads
    from sawsim_amd.plan import kernel_work
    cells = workloads.config5(n_cells=args.cells)
    t0 = time.time()
    engines = []
    for i, c in enumerate(cells):
        engines.append(build_engine(c))
        torch.cuda.synchronize()
        log(f'[bench] cell {i}: {c.model.ne} elements, n={engines[-1].n_active}, {time.time() - t0:.0f}s')
    setup = time.time() - t0
    P = cells[0].points
    from sawsim_amd.sweep import check_workspace, hbm_available
    torch.cuda.empty_cache()          # the cells' assembly temporaries
    ne = [c.model.ne for c in cells]
    if args.share:
        for eng in engines:
            eng.set_share(args.share)
    tuning = None
    if not args.no_autotune:
        # kernel knobs and factorization sharing tuned on the largest cell, applied to every cell
        from sawsim_amd.tune import autotune, tune_share, _SETTERS
        big = int(np.argmax(ne))
        fb = cells[big].frequencies()
        tuning = autotune(engines[big], fb, 'mixed', gradient=True)
        if not args.share:
            tuning['share'] = tune_share(engines[big], fb, 'mixed', gradient=True)
        for eng in engines:
            for knob, v in tuning['choice'].items():
                getattr(eng, _SETTERS[knob])(v)
            eng.set_share(engines[big].share)
        engines[big].release_workspace()
        log(f"[bench] autotune on cell {big}: {tuning['choice']}, share {engines[big].share} "
            f"({tuning.get('share', {}).get('ms')} ms); memory: {_mem()}")
    budget = args.arena_gb * 1e9 if args.arena_gb else hbm_available()
    batches = []
    for eng in engines:
        b = min(P, eng.auto_batch('mixed', gradient=True, budget=budget))
        b = -(-P // -(-P // b))
        batches.append(b)
    # every cell's chunk has to fit beside the shared workspace, which the largest arena sets
    ws_bytes = max(eng.chunk_arena_bytes(b, 'mixed') for eng, b in zip(engines, batches))
    extra = [eng.chunk_bytes(b, 'mixed', True) - eng.chunk_arena_bytes(b, 'mixed') for eng, b in zip(engines, batches)]
    while ws_bytes + max(extra) > budget:
        i = int(np.argmax([eng.chunk_arena_bytes(b, 'mixed') + x for eng, b, x in zip(engines, batches, extra)]))
        if batches[i] == 1:
            break
        batches[i] = -(-P // (-(-P // batches[i]) + 1))       # one more chunk for that cell
        extra[i] = engines[i].chunk_bytes(batches[i], 'mixed', True) - engines[i].chunk_arena_bytes(batches[i], 'mixed')
        ws_bytes = max(eng.chunk_arena_bytes(b, 'mixed') for eng, b in zip(engines, batches))
    check_workspace(ws_bytes + max(extra), what='shared front workspace and chunk vectors')
    log(f'[bench] memory: {_mem()}; budget {budget / 1e9:.1f} GB, largest chunk {max(extra) / 1e9:.1f} GB '
        f'beside the workspace')
    shared = torch.empty(ws_bytes, dtype=torch.uint8, device='cuda')
    for eng in engines:
        eng._ws = shared                       # every cell reuses the one front workspace
    log(f'[bench] config5: {len(cells)} cells, {sum(ne)} elements ({min(ne)}-{max(ne)}), setup {setup:.0f}s, '
        f'workspace {ws_bytes / 1e9:.0f} GB, share {engines[0].share}, chunks {[-(-P // b) for b in batches]}')

    def sweep(i):
        return engines[i].sweep_charges(cells[i].frequencies(), precision='mixed', batch=batches[i], gradient=True)
    for _ in range(args.warmup):
        sweep(int(np.argmax(ne)))
    torch.cuda.synchronize()
    for eng in engines:
        _lib.call('sawfem_mf_profile', eng._plan, 1)
    lat, results = [], []
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    for i in range(len(cells)):
        t1 = time.perf_counter()
        results.append(sweep(i))
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    dt = time.perf_counter() - t0
    log(f'[bench] peak torch allocation over the sweeps {torch.cuda.max_memory_allocated() / 1e9:.1f} GB; {_mem()}')
    schur_ms = schur_n = 0.0
    schur_flops = 0.0
    for eng, c, res in zip(engines, cells, results):
        prof, _ = profile_read(eng)
        schur_ms += prof['gemm_s'][0]
        schur_n += prof['gemm_s'][1]
        schur_flops += kernel_work(eng.mf, 64)['schur_flops'] * sum(it['factorizations'] for it in res[2])
        _lib.call('sawfem_mf_profile', eng._plan, 0)
    npts = sum(c.points for c in cells)
    value = npts / dt
    lat_ms = np.array(lat) * 1e3
    log(f'[bench] design set: {dt:.2f}s for {npts} points -> {value:.1f} pts/s; latency p50 '
        f'{np.percentile(lat_ms, 50):.0f} ms p90 {np.percentile(lat_ms, 90):.0f} ms')
    # accuracy against the FP64 path on every cell, at every 8th of its points, on one shared
    # workspace sized for the largest cell's FP64 chunk
    k = np.arange(0, P, 8)
    for eng in engines:
        eng._ws = None
    del shared
    torch.cuda.empty_cache()
    fp64_batch = [min(k.size, eng.auto_batch('fp64', gradient=True)) for eng in engines]
    shared = torch.empty(max(eng.arena_bytes(b, 'fp64') for eng, b in zip(engines, fp64_batch)), dtype=torch.uint8,
                         device='cuda')
    for eng in engines:
        eng._ws = shared
    errs, gerrs = [], []
    t1 = time.perf_counter()
    for i in range(len(cells)):
        Q, dY, _ = results[i]
        Q64, dY64, _ = engines[i].sweep_charges(cells[i].frequencies()[k], precision='fp64', gradient=True,
                                                batch=fp64_batch[i])
        errs.append(float(np.max(np.abs(Q[k] - Q64) / np.abs(Q64))))
        scale = np.abs(cells[i].params)
        scale[12:15] = 1.0
        gerrs.append(float(np.max(np.abs((dY[k] - dY64) * scale).max(axis=1) / np.abs(dY64 * scale).max(axis=1))))
    log(f'[bench] mixed vs fp64 on {len(errs)} cells x {k.size} points ({time.perf_counter() - t1:.0f}s): '
        f'Y11 {max(errs):.2e}, dY/dp {max(gerrs):.2e}')
    # the plain FP64 path on the median-size cell, and the mixed rate on that same cell
    mid = int(np.argsort(ne)[len(ne) // 2])
    plain = plain_adj = plain_fd = None
    if args.plain_points > 0:
        plain, plain_adj, plain_fd = plain_fp64(engines[mid], cells[mid], cells[mid].frequencies(),
                                                min(args.plain_points, P))
    mid_rate = cells[mid].points / lat[mid]
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(cells[mid], args.cpu_workers)
        except Exception as exc:
            log(f'[bench] cpu baseline failed: {exc!r}')
    achieved = schur_flops / (schur_ms * 1e-3) / 1e12 if schur_ms > 0 else None
    out = dict(
        metric='freq pts/s (Y11 + dY11/dp) at Y11 rel err <=1e-6 vs FP64; vs plain FP64 GPU',
        value=value, unit='freq pts/s', n_gpus=1, steps=1, warmup=args.warmup, ms_per_step=1e3 * dt,
        higher_is_better=True, scaling='weak', vs_baseline=None,
        dtype='complex64 fronts (fp32 MFMA) + fp64 refinement', data='synthetic',
        config=dict(workload=f'config5: {len(cells)} design cells (seeds 0..{len(cells) - 1}), {sum(ne)} Q9 elements '
                             f'({min(ne)}-{max(ne)} per cell), {P} points each, swept back to back',
                    cells=len(cells), points_per_cell=P, elements=ne,
                    n_active=[int(e.n_active) for e in engines], batches=batches, tuning=tuning,
                    share=engines[0].share,
                    factorizations=int(sum(it['factorizations'] for r in results for it in r[2])),
                    gradient='dY11/dp for 17 parameters by adjoint'),
        latency_ms_per_cell=dict(p50=float(np.percentile(lat_ms, 50)), p90=float(np.percentile(lat_ms, 90)),
                                 p99=float(np.percentile(lat_ms, 99)), max=float(lat_ms.max()),
                                 min=float(lat_ms.min())),
        per_cell_ms=[round(float(x), 1) for x in lat_ms],
        y11_max_rel_err_vs_fp64=max(errs), dy11_max_rel_err_vs_fp64_adjoint=max(gerrs),
        accuracy_check='every cell at every 8th of its points against the FP64 path with its FP64 adjoint',
        y11_max_rel_err_per_cell=[float(f'{e:.3g}') for e in errs],
        plain_fp64_gpu_pts_per_s=dict(cell=mid, elements=ne[mid], y11_only=plain, y11_plus_fp64_adjoint=plain_adj,
                                      y11_plus_central_fd=plain_fd, mixed_on_that_cell=mid_rate)
        if plain else None,
        speedup_vs_plain_fp64=dict(central_fd=mid_rate / plain_fd, fp64_adjoint=mid_rate / plain_adj,
                                   y11_only=mid_rate / plain) if plain else None,
        roofline=dict(bound='mfma', achieved=achieved, peak=PEAK_FP32_MFMA, unit='TFLOP/s',
                      frac=(achieved / PEAK_FP32_MFMA) if achieved else None, traffic=None,
                      kernel='k_gemm<float, 1> (Schur complement) over all cells', launches=int(schur_n),
                      kernel_ms=schur_ms, algorithmic_flops=schur_flops),
        setup_s=setup,
        setup_s_per_cell=setup / len(cells),
        end_to_end_pts_per_s=npts / (setup + dt),
        end_to_end_note='cells built one after another (assembly, Bloch layout, symbolic analysis, plan upload) '
                        'then swept; kernel tuning and the workspace allocation excluded',
        cpu_baseline=cpu,
    )
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='config2')
    ap.add_argument('--batch', type=int, default=0, help='frequencies per factorization chunk (0: fit --arena-gb)')
    ap.add_argument('--arena-gb', type=float, default=0.0,
                    help='device memory per chunk when --batch is 0 (0: what is free above the reserve)')
    ap.add_argument('--cells', type=int, default=64, help='design-set size (config5)')
    ap.add_argument('--plain-points', type=int, default=8)
    ap.add_argument('--check-points', type=int, default=0,
                    help='frequencies of the FP64 accuracy check (0: every point of the sweep)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-workers', type=int, default=8)
    ap.add_argument('--leaf-weight', type=int, default=0, help='nested-dissection leaf size (0: default)')
    ap.add_argument('--small-max', type=int, default=-1, help='largest front for the fused kernel (-1: default)')
    ap.add_argument('--solve-full', type=int, default=-1,
                    help='largest front solved by one workgroup (-1: the library default)')
    ap.add_argument('--panel-lu', type=int, default=-1,
                    help='largest front order whose 32-wide panels use the one-kernel panel LU (-1: default)')
    ap.add_argument('--tri-lds', type=int, default=-1,
                    help='LDS bytes for the factor block of the LDS-staged solves (-1: default, 0: off)')
    ap.add_argument('--share', type=int, default=0,
                    help='frequencies per complex64 factorization (0: chosen by the tuner, 1 without it)')
    ap.add_argument('--no-autotune', action='store_true',
                    help='keep the library defaults of the kernel-selection knobs (no tuning run)')
    ap.add_argument('--copy-mib', type=int, default=0,
                    help='only run a plain device copy of this many MiB (HBM counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_mib:
        import torch
        n = args.copy_mib * 2 ** 20 // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes_read=4 * n, copy_bytes_written=4 * n)), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps(dict(error='this build runs on exactly one GPU; multi-GPU sharding of the '
                                    'frequency axis is designed but not built', n_gpus=args.gpus)))
        sys.exit(2)
    import torch
    from sawsim_amd import _lib, workloads
    from sawsim_amd.plan import kernel_work
    assert torch.cuda.is_available()
    if args.workload == 'config5':
        return design_set(args)
    cell = workloads.WORKLOADS[args.workload]()
    hex_cell = isinstance(cell, workloads.HexCell)
    freqs = cell.frequencies()
    P = freqs.size
    t0 = time.time()
    eng = build_engine(cell, **({'leaf_weight': args.leaf_weight} if args.leaf_weight else {}))
    if args.small_max >= 0:
        _lib.call_value('sawfem_mf_set_small', eng._plan, 0, args.small_max)
    if args.solve_full >= 0:
        eng.set_solve_full(args.solve_full)
    if args.panel_lu >= 0:
        eng.set_panel_lu(args.panel_lu)
    if args.tri_lds >= 0:
        eng.set_tri_lds(args.tri_lds)
    torch.cuda.synchronize()
    setup = time.time() - t0
    log(f'[bench] {args.workload}: n={eng.n_active} fronts={eng.mf.nf} setup {setup:.2f}s '
        f'(analysis {eng.analysis_seconds:.2f}s)')
    asm = assembly_roofline(eng, cell)
    log('[bench] assembly: ' + ', '.join(f"{k} {v['achieved_gbs']:.0f} GB/s ({v['ms_per_launch']:.2f} ms)"
                                         for k, v in asm.items()))
    nb = 64                                   # panel width of the complex64 factorization
    work = kernel_work(eng.mf, nb)
    budget = args.arena_gb * 1e9 if args.arena_gb else None

    def chunking():
        b = args.batch or min(P, eng.auto_batch('mixed', gradient=True, budget=budget))
        return -(-P // -(-P // b)), -(-P // b)

    if args.share:
        eng.set_share(args.share)
    tuning = None
    if not args.no_autotune:
        from sawsim_amd.tune import autotune, tune_share
        t0 = time.time()
        tuning = autotune(eng, freqs, 'mixed', gradient=True, batch=chunking()[0])
        log(f"[bench] autotune {time.time() - t0:.1f}s: {tuning['choice']} "
            f"({ {k: {c: round(v, 1) for c, v in d.items()} for k, d in tuning['ms'].items()} } ms)")
        if not args.share:
            t0 = time.time()
            tuning['share'] = tune_share(eng, freqs, 'mixed', gradient=True)
            log(f"[bench] factorization sharing {time.time() - t0:.1f}s: share {tuning['share']['choice']} "
                f"(ms per frequency { {c: round(v, 2) for c, v in tuning['share']['ms_per_frequency'].items()} } "
                f"over chunks of {tuning['share']['batch']} frequencies)")
    batch, nchunk = chunking()
    from sawsim_amd.sweep import hbm_available
    log(f'[bench] share {eng.share}: batch {batch} x {nchunk} chunks, arena '
        f'{eng.chunk_arena_bytes(batch, "mixed") / 1e9:.1f} GB, chunk '
        f'{eng.chunk_bytes(batch, "mixed", True) / 1e9:.1f} GB of '
        f'{hbm_available(0 if getattr(eng, "_ws", None) is None else eng._ws.numel()) / 1e9:.1f} GB available')

    def sweep():
        return eng.sweep_charges(freqs, precision='mixed', batch=batch, gradient=True)
    for _ in range(args.warmup):
        sweep()
    torch.cuda.synchronize()
    # timed region: factorizations as replayed HIP graphs; HIP events around the grouped solve
    # kernel's launches only (profile mode 2: the roofline input).  Without sharing no grouped
    # solves run: every launch is timed instead (mode 1, factorizations launch by launch) and the
    # Schur GEMM is the roofline entry.
    mode = 2 if eng.share > 1 else 1
    _lib.call('sawfem_mf_profile', eng._plan, mode)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Q, dY, info = sweep()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    import ctypes as C
    g_ms, g_bytes, g_n = C.c_double(), C.c_double(), C.c_int64()
    _lib.call('sawfem_mf_profile_grp', eng._plan, C.byref(g_ms), C.byref(g_bytes), C.byref(g_n))
    grp = dict(ms=g_ms.value, bytes=g_bytes.value, launches=g_n.value)
    if mode == 1:
        prof, prof_levels = profile_read(eng)
    _lib.call('sawfem_mf_profile', eng._plan, 0)
    value = args.steps * P / dt
    log(f'[bench] mixed sweep + adjoint {dt / args.steps:.3f}s/step -> {value:.1f} pts/s; ir {info}')
    nprof, eager = args.steps, None
    if mode == 2:
        # per-kernel breakdown: NPROF more sweeps with every launch timed (profile mode 1; the
        # factorizations then run launch by launch, not as graphs)
        nprof = 1
        _lib.call('sawfem_mf_profile', eng._plan, 1)
        t0 = time.perf_counter()
        for _ in range(nprof):
            sweep()
        torch.cuda.synchronize()
        eager = dict(pts_per_s=nprof * P / (time.perf_counter() - t0),
                     nodes_per_factorization=int(_lib.call_value('sawfem_mf_graph_nodes', eng._plan)),
                     note=f'{nprof} sweep(s) after the timed region with every launch timed by HIP events and the '
                          'factorizations launch by launch (the timed region replays each factorization as one '
                          'captured HIP graph of nodes_per_factorization kernels)')
        prof, prof_levels = profile_read(eng)
        _lib.call('sawfem_mf_profile', eng._plan, 0)
        log(f'[bench] profiled sweep, eager factorizations: {eager["pts_per_s"]:.1f} pts/s; '
            f'{eager["nodes_per_factorization"]} graph nodes per factorization in the timed region')
    # accuracy of the mixed path against the plain FP64 path (with its FP64 adjoint) on a sample
    torch.cuda.empty_cache()          # the FP64 chunks have another size than the timed ones
    # accuracy of the mixed path against the FP64 path (batched complex128 fronts, FP64 adjoint)
    idx = np.arange(P) if not args.check_points else np.linspace(0, P - 1, min(args.check_points, P)).astype(int)
    t0 = time.perf_counter()
    Q64, dY64, _ = eng.sweep_charges(freqs[idx], precision='fp64', gradient=True)
    log(f'[bench] FP64 check sweep of {len(idx)} points: {time.perf_counter() - t0:.1f}s')
    y_err_pts = np.abs(Q[idx] - Q64) / np.abs(Q64)
    y_err = float(np.max(y_err_pts))
    frfa = None
    if len(idx) == P:
        from sawsim_amd.metrics import compare, resonances
        sign = 1j if hex_cell else -1j           # Y = +i w Q / V on the slabs, -i w Q / V in 2D
        ym, y64 = sign * 2 * np.pi * freqs * Q, sign * 2 * np.pi * freqs * Q64
        rm, r64 = resonances(freqs, ym), resonances(freqs, y64)
        d = compare(r64, rm)
        frfa = dict(fp64_fr_hz=r64['fr_hz'], fp64_fa_hz=r64['fa_hz'], mixed_fr_hz=rm['fr_hz'], mixed_fa_hz=rm['fa_hz'],
                    fr_diff_hz=d['fr_diff_hz'], fa_diff_hz=d['fa_diff_hz'], step_hz=r64['step_hz'],
                    warnings=r64['warnings'])
    scale = np.abs(cell.params)
    scale[12:15] = 1.0
    g_err = float(np.max(np.abs((dY[idx] - dY64) * scale).max(axis=1) / np.abs(dY64 * scale).max(axis=1)))
    # plain FP64 GPU path: one frequency per factorization
    pp = min(args.plain_points, P)
    plain = plain_adj = plain_fd = None
    if pp > 0:
        plain, plain_adj, plain_fd = plain_fp64(eng, cell, freqs, pp)
    if pp > 0:
        log(f'[bench] plain fp64: Y11 {plain:.2f} pts/s, +adjoint {plain_adj:.2f}, +central FD {plain_fd:.3f}')
    log(f'[bench] mixed vs fp64: Y11 {y_err:.2e}, dY/dp {g_err:.2e}')
    # dominant MFMA kernel: the Schur-complement GEMM k_gemm<float, 1> (one launch per level)
    schur_ms, schur_n = prof['gemm_s']
    # factorizations per sweep: one per `share` frequencies, plus the complex64 refactorizations
    # of frequencies whose shared factors did not converge
    nfact = sum(it['factorizations'] for it in info)
    schur_flops = work['schur_flops'] * nfact * nprof
    achieved = schur_flops / (schur_ms * 1e-3) / 1e12 if schur_ms > 0 else None
    schur_alg = work['schur_bytes'] * nfact * nprof / max(schur_n, 1)
    schur_traffic, schur_traffic_note = counter_traffic('schur', args.workload, eng.share, schur_alg)
    lupd_tf = work['lupd_flops'] * nfact * nprof / (prof['gemm_l'][0] * 1e-3) / 1e12 if prof['gemm_l'][0] else None
    # solves: each forward+backward pass streams the stored L/U entries (complex64) once
    # (passes over the factors: paired forward+adjoint solves share each pass)
    nsolve_pts = sum(sum(it['solved_per_pass']) if it.get('paired_adjoint')
                     else it['points'] * (2 + it['ir_steps'] + it.get('adj_steps', 0)) for it in info) * nprof
    solve_bytes = nsolve_pts * work['lu_entries'] * 8
    solve_ms = prof['fwd'][0] + prof['bwd'][0]
    solve_gbs = solve_bytes / (solve_ms * 1e-3) / 1e9 if solve_ms else None
    # whole sweep: algorithmic flops of factorizations, solves (every right-hand side) and FP64
    # residuals, over the wall time of the timed region
    nnz_aa = int(eng.bloch.nnz_aa)
    def chunk_flops(it):
        if it.get('paired_adjoint'):       # two right-hand sides per solved frequency and pass
            return (it['factorizations'] * work['factor_flops'] + 2 * sum(it['solved_per_pass']) * work['solve_flops']
                    + 2 * sum(it['solved_per_pass'][1:]) * 8 * nnz_aa)
        return (it['factorizations'] * work['factor_flops']
                + it['points'] * ((2 + it['ir_steps'] + it.get('adj_steps', 0)) * work['solve_flops']
                                  + (it['ir_steps'] + it.get('adj_steps', 0)) * 8 * nnz_aa))
    sweep_flops = args.steps * sum(chunk_flops(it) for it in info)
    sweep_tf = sweep_flops / dt / 1e12
    mfma_roofline_entry = dict(bound='mfma', achieved=achieved, peak=PEAK_FP32_MFMA, unit='TFLOP/s',
                               frac=(achieved / PEAK_FP32_MFMA) if achieved else None,
                               traffic=schur_traffic, traffic_note=schur_traffic_note,
                               kernel='k_gemm<float, 1, 64, 64, *> (Schur complement, complex64 on FP32 MFMA)',
                               launches=schur_n, kernel_ms=schur_ms, algorithmic_flops=schur_flops,
                               algorithmic_bytes_per_launch=schur_alg, avg_launch_ms=schur_ms / max(schur_n, 1),
                               note=(f'from the profiled sweep after the timed region ({nprof} sweep)' if eager
                                     else 'from the timed region (every launch timed)'))
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(cell, args.cpu_workers)
        except Exception as exc:  # the baseline is reported, never required
            log(f'[bench] cpu baseline failed: {exc!r}')
    out = dict(
        metric='freq pts/s (Y11 + dY11/dp) at Y11 rel err <=1e-6 vs FP64; vs plain FP64 GPU',
        value=value, unit='freq pts/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * dt / args.steps, higher_is_better=True, scaling='weak', vs_baseline=None,
        dtype='complex64 fronts (fp32 MFMA) + fp64 refinement', data='synthetic',
        config=dict(workload=f'{args.workload}: {cell.mesh.ne if hex_cell else cell.model.ne} '
                             f'{"Hex27" if hex_cell else "Q9"} elements, {eng.n_active} complex unknowns, '
                             f'{P} frequency points, batch {batch}',
                    points=P, n_active=eng.n_active, fronts=eng.mf.nf, batch=batch,
                    gradient=f'dY11/dp for {eng.n_params} parameters (3m constants, Euler angles, beta_dK, '
                             f'tan delta{", rho" if hex_cell else ""}) by adjoint',
                    tuning=tuning, share=eng.share, factorizations_per_sweep=nfact),
        y11_max_rel_err_vs_fp64=y_err,
        y11_check_points=int(len(idx)),
        y11_rel_err_per_point=dict(freq_hz=[float(f) for f in freqs[idx]],
                                   rel_err=[float(f'{e:.3g}') for e in y_err_pts]),
        setup_s=setup,
        end_to_end_pts_per_s=P / (setup + dt / args.steps),
        end_to_end_note='one sweep from the cell description: assembly, Bloch layout, symbolic analysis, '
                        'plan upload and the sweep (kernel tuning excluded)',
        fr_fa_vs_fp64=frfa,
        dy11_max_rel_err_vs_fp64_adjoint=g_err,
        plain_fp64_gpu_pts_per_s=dict(y11_only=plain, y11_plus_fp64_adjoint=plain_adj,
                                      y11_plus_central_fd=plain_fd) if pp > 0 else None,
        speedup_vs_plain_fp64=dict(central_fd=value / plain_fd, fp64_adjoint=value / plain_adj,
                                   y11_only=value / plain) if pp > 0 else None,
        ir=info,
        roofline=grp_roofline(grp, args.workload, eng.share) or dict(
            mfma_roofline_entry, note='no grouped solves ran (share 1): the Schur GEMM, timed in the timed region, '
                                      'is the roofline entry'),
        mfma_roofline=mfma_roofline_entry,
        other_kernels=dict(
            whole_sweep=dict(bound='mfma', achieved=sweep_tf, peak=PEAK_FP32_MFMA, unit='TFLOP/s',
                             frac=sweep_tf / PEAK_FP32_MFMA, algorithmic_flops=sweep_flops,
                             note='(factor + every solve + FP64 residuals) / wall time of the timed sweep'),
            lupdate_gemm_tflops=lupd_tf,
            solve=dict(bound='hbm', achieved_gbs=solve_gbs, peak_gbs=8000.0,
                       frac=(solve_gbs / 8000.0) if solve_gbs else None, bytes=solve_bytes, ms=solve_ms,
                       note='forward+backward passes; algorithmic bytes = stored L/U entries x 8 B per point; '
                            'ms = summed HIP-event spans, which overlap across the size buckets of a level')),
        assembly_roofline=dict(kernels=asm, bound='hbm', peak_gbs=8000.0,
                               note='K1 element matrices, K2 node-pair gather into K/M/D, K3 Bloch gathers; '
                                    'algorithmic bytes over HIP-event time, mean of 3 re-assemblies'),
        kernel_ms={k: v[0] for k, v in prof.items()},
        kernel_ms_note='per kind, summed HIP-event spans of its launches; the size buckets of a level run their '
                       'panel kernels and solves concurrently on side streams, so those sums exceed their wall '
                       'time; the Schur complements run one at a time on the launch stream',
        kernel_ms_per_level=prof_levels,
        factor_flops_per_point=work['factor_flops'],
        profiled_sweep=eager,
        cpu_baseline=cpu,
    )
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
