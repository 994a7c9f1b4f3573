"""GPU: the pivot guard on a planted zero pivot, and the complex128 refactorization that takes
over the frequencies whose complex64 refinement does not converge."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from sawsim_amd import workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _host_system(eng, w, V=1.0):
    """The engine's active system at w from its device arrays (A, rhs, kts_a, kts_z)."""
    bl = eng.bloch
    n = eng.n_active
    Kaa, Maa = eng.Kaa.cpu().numpy(), eng.Maa.cpu().numpy()
    A = Kaa - w * w * Maa
    rhs = -(eng.rK.cpu().numpy() - w * w * eng.rM.cpu().numpy()) * V
    if eng.Daa is not None:
        A = A + 1j * w * eng.beta * eng.Daa.cpu().numpy()
        rhs = rhs - 1j * w * eng.beta * eng.rD.cpu().numpy() * V
    A = sp.csr_matrix((A, bl.aa_indices, bl.aa_indptr), shape=(n, n))
    return A, rhs, eng.kts_a.cpu().numpy(), eng.kts_z


def _refined(A, rhs, steps=4):
    lu = spla.splu(A.tocsc())
    x = lu.solve(rhs)
    for _ in range(steps):
        x = x + lu.solve(rhs - A @ x)
    return x


def test_planted_zero_pivot_is_guarded_and_the_sweep_stays_exact():
    """K_gg = w^2 M_gg on the first pivot row of a childless front: without pivoting the
    factorization would divide by (almost exactly) zero there.  The guard replaces that pivot
    (counted in both precisions); the mixed sweep then either refines or refactors, and Q
    equals the refined SuperLU solution of the modified system."""
    _say('[test_planted_zero_pivot_is_guarded_and_the_sweep_stays_exact]')
    cell = workloads.double_layer(points=3)
    eng = workloads.build_engine(cell)
    mf = eng.mf
    leaf = int(mf.lvl_fronts[0])
    g = int(mf.rows[mf.rows_ptr[leaf]])
    bl = eng.bloch
    r = np.repeat(np.arange(eng.n_active), np.diff(bl.aa_indptr))
    dslot = int(np.flatnonzero((r == g) & (bl.aa_indices == g))[0])
    w = 2 * np.pi * 2.0e9
    eng.Kaa[dslot] = w * w * eng.Maa[dslot]
    A, rhs, kts_a, kts_z = _host_system(eng, w)
    Qref = kts_a @ _refined(A, rhs) + kts_z
    for precision in ('fp64', 'mixed'):
        res = eng.solve_batch([w], 1.0, precision=precision)
        err = abs(res['Q'][0] - Qref) / abs(Qref)
        _say(precision, 'perturbed', res['perturbed'], res.get('perturbed_fp64'), 'refactored', res['refactored'],
             'min ratio', res['pivot_min_ratio'], 'err', err)
        assert res['perturbed'] >= 1
        if res['refactored'][0]:
            assert res['perturbed_fp64'] >= 1
        assert err <= 1e-8


@pytest.mark.parametrize('gradient', [False, True])
def test_stalled_refinement_is_refactored_in_fp64(gradient):
    """max_ir = 1 leaves every complex64 frequency above ir_tol: all of them go to the
    complex128 refactorization (in chunks of half the workspace), and Q, dY and lambda equal
    the FP64 path's bit for bit.  A normal run refactors none and replaces no pivot."""
    _say(f'[test_stalled_refinement_is_refactored_in_fp64] gradient={gradient}')
    cell = workloads.config3(mesh_um=0.12, points=5)
    eng = workloads.build_engine(cell, gradient=True)
    om = 2 * np.pi * cell.frequencies()
    ref = eng.solve_batch(om, precision='fp64', max_ir=1, gradient=gradient)
    got = eng.solve_batch(om, precision='mixed', max_ir=1, gradient=gradient)
    _say('refactored', got['refactored'], 'perturbed', got['perturbed'], got.get('perturbed_fp64'))
    assert got['refactored'].all()
    assert np.array_equal(got['Q'], ref['Q'])
    assert np.array_equal(got['x'].cpu().numpy(), ref['x'].cpu().numpy())
    if gradient:
        assert np.array_equal(got['dY'], ref['dY'])
        assert np.array_equal(got['lam'].cpu().numpy(), ref['lam'].cpu().numpy())
    normal = eng.solve_batch(om, precision='mixed', gradient=gradient)
    assert not normal['refactored'].any()
    assert normal['perturbed'] == 0 and 0 < normal['pivot_min_ratio'] < 1
    assert np.max(np.abs(normal['Q'] - ref['Q']) / np.abs(ref['Q'])) <= 1e-6
