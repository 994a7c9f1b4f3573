"""GPU: the refined chunk sweep of the C ABI (sawfem_sweep_chunk) against the Python engine's
solve_batch on the same kernels, and its status codes (refactorization used, non-finite charge,
workspace too small)."""
import ctypes as C
import sys

import numpy as np
import pytest

from sawsim_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_native_chunk_equals_solve_batch(precision):
    _say(f'[test_native_chunk_equals_solve_batch] {precision}')
    cell = workloads.config3(mesh_um=0.09, points=7)
    eng = workloads.build_engine(cell)
    om = 2 * np.pi * cell.frequencies()
    ref = eng.solve_batch(om, precision=precision)
    Q, err, flags, rc = eng.sweep_chunk_native(om, precision=precision)
    _say(rc, flags, np.abs(Q - ref['Q']).max() / np.abs(ref['Q']).max(), err)
    assert rc == _lib.OK and not flags.any()
    # the same kernels and refinement decisions: Q bit for bit; the error estimates are host
    # arithmetic in another order (last bits)
    assert np.array_equal(Q, ref['Q']) and np.allclose(err, ref['err'], rtol=1e-12, atol=0)


def test_native_chunk_status_codes():
    _say('[test_native_chunk_status_codes]')
    cell = workloads.config3(mesh_um=0.09, points=5)
    eng = workloads.build_engine(cell)
    om = 2 * np.pi * cell.frequencies()
    # one correction leaves every complex64 frequency above ir_tol: all refactored (status 6)
    Q, err, flags, rc = eng.sweep_chunk_native(om, max_ir=1)
    ref = eng.solve_batch(om, precision='fp64', max_ir=1)
    assert rc == _lib.REFACTORED and np.all(flags & 1) and np.array_equal(Q, ref['Q'])
    # workspace too small: refused before any launch (status 4)
    prec, B = 0, om.size
    op = _lib.SweepOperator(eng.n_active, eng.aa_indptr.data_ptr(), eng.aa_indices.data_ptr(), eng.Kaa.data_ptr(),
                            eng.Maa.data_ptr(), eng._dptr(eng.Daa), eng.rK.data_ptr(), eng.rM.data_ptr(),
                            eng._dptr(eng.rD), eng.kts_a.data_ptr(), eng.kts_z.real, eng.kts_z.imag, eng.beta)
    need = _lib.load().sawfem_sweep_chunk_bytes(eng._plan, eng.n_active, B, prec)
    ws = eng._workspace(need)
    out = np.empty(2 * B), np.empty(B), np.zeros(B, dtype=np.int32)
    from sawsim_amd.piezo_fem import stream_ptr
    omc = np.ascontiguousarray(om)
    rc = _lib.raw('sawfem_sweep_chunk', eng._plan, stream_ptr(), C.byref(op), B, omc.ctypes.data, 1.0, prec, 1e-8, 8,
                  1, ws.data_ptr(), int(need) - 1, out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data)
    assert rc == _lib.OUT_OF_MEMORY
    rc = _lib.raw('sawfem_sweep_chunk', eng._plan, stream_ptr(), C.byref(op), B, omc.ctypes.data, -1.0, prec, 1e-8, 8,
                  1, ws.data_ptr(), int(need), out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data)
    assert rc == _lib.INVALID
    # a NaN in the operator gives a non-finite charge: status 5 (the reference raises ValueError)
    eng.Kaa[0] = complex(np.nan, 0.0)
    Q, err, flags, rc = eng.sweep_chunk_native(om)
    assert rc == _lib.NON_FINITE and not np.all(np.isfinite(Q))


@pytest.mark.parametrize('max_ir', [8, 1])
def test_native_shared_chunk_equals_solve_batch(max_ir):
    """sawfem_sweep_chunk_share against solve_batch(share=3) on 7 frequencies spread over the band
    (groups of 3, 3 and a partial group of 1): Q bit for bit and the same fallbacks (the
    frequencies far from their group's mean take their own complex64 factors; with one
    correction allowed they go on to complex128)."""
    _say(f'[test_native_shared_chunk_equals_solve_batch] max_ir={max_ir}')
    cell = workloads.config3(mesh_um=0.09, points=7)
    eng = workloads.build_engine(cell)
    om = 2 * np.pi * cell.frequencies()
    ref = eng.solve_batch(om, precision='mixed', share=3, max_ir=max_ir)
    Q, err, flags, rc = eng.sweep_chunk_native(om, share=3, max_ir=max_ir)
    _say(rc, flags, ref['factorizations'], ref['own_factors'], ref['refactored'])
    assert np.array_equal(Q, ref['Q'])
    assert np.array_equal((flags & 4) != 0, ref['own_factors'])
    assert np.array_equal((flags & 1) != 0, ref['refactored'])
    assert ref['factorizations'] == 3 + int(ref['own_factors'].sum())
    assert rc == (_lib.REFACTORED if ref['refactored'].any() else _lib.OK)
    if max_ir == 1:
        assert ref['refactored'].all()
