"""GPU adjoint dY11/dp against central finite differences (GPU FP64 path and the reference)."""
import sys

import numpy as np
import pytest

from sawsim_amd import params as prm, workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fd_gpu(cell, eng, f, p0, V=1.0):
    """Central differences of the FP64 GPU sweep, reusing the symbolic plan."""
    out = np.zeros((17, f.size), dtype=complex)
    for k in range(17):
        h = 1e-6 * abs(p0[k]) if p0[k] != 0 else 1e-6 * 180 / np.pi
        ys = []
        for sgn in (1, -1):
            p = p0.copy()
            p[k] += sgn * h
            K, M, D = workloads.assemble_cell(cell, workloads.materials_for_params(cell, p))
            eng.update_operator(K, M, D, beta=p[15])
            Q, _ = eng.sweep_charges(f, precision='fp64')
            ys.append(-1j * 2 * np.pi * f * Q / V)
        out[k] = (ys[0] - ys[1]) / (2 * h)
    K, M, D = workloads.assemble_cell(cell)
    eng.update_operator(K, M, D, beta=p0[15])
    return out


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_adjoint_matches_gpu_finite_differences(precision):
    _say(f'[test_adjoint_matches_gpu_finite_differences] {precision}')
    cell = workloads.config3(mesh_um=0.12, points=3)
    eng = workloads.build_engine(cell, gradient=True)
    f = cell.frequencies(3)
    Q, dY, info = eng.sweep_charges(f, precision=precision, gradient=True)
    fd = _fd_gpu(cell, eng, f, cell.params.copy())
    scale = np.abs(cell.params)[:, None]
    scale[12:15] = 1.0
    a, b = dY.T * scale, fd * scale
    err = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
    _say(precision, 'adjoint vs FD, per frequency', err, info)
    assert err.max() <= 1e-5


def test_adjoint_matches_reference_finite_differences():
    """dY11/dp vs central differences of the reference sweep (refined SuperLU) on config 3."""
    _say('[test_adjoint_matches_reference_finite_differences]')
    from golden_cells import cfg3_coarse_cell
    cell, g = cfg3_coarse_cell()
    eng = workloads.build_engine(cell, gradient=True)
    f = g['frequency_hz']
    Q, dY, info = eng.sweep_charges(f, precision='mixed', gradient=True)
    y = -1j * 2 * np.pi * f * Q
    ey = np.abs(y - g['y11_refined']) / np.abs(g['y11_refined'])
    scale = np.abs(cell.params)[:, None]
    scale[12:15] = 1.0
    a, b = dY.T * scale, g['dy11_refined'] * scale
    err = np.abs(a - b).max(axis=0) / np.abs(b).max(axis=0)
    _say('Y11 vs reference', ey, 'dY vs reference FD', err)
    assert ey.max() <= 1e-8
    assert err.max() <= 1e-4


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_paired_adjoint_matches_transposed_solves(precision):
    """The adjoint as a second right-hand side (lambda = D A^-1 D^-1 g) against transposed
    solves with the same factors, on config 3 (damping, dielectric loss, PML).  Both are refined
    to ir_tol 1e-12, so that the comparison tests the algebra rather than where each path's
    stopping rule lands (at the default 1e-8 the paired path may stop on its dual-weighted
    residual one step earlier)."""
    _say(f'[test_paired_adjoint_matches_transposed_solves] {precision}')
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell, gradient=True)
    assert eng.scaled_symmetric()
    om = 2 * np.pi * cell.frequencies(4)
    a = eng.solve_batch(om, precision=precision, gradient=True, adjoint='paired', ir_tol=1e-12, max_ir=12)
    b = eng.solve_batch(om, precision=precision, gradient=True, adjoint='transposed', ir_tol=1e-12, max_ir=12)
    assert a.get('paired') and not b.get('paired')
    lam_err = float(((a['lam'] - b['lam']).abs().max() / b['lam'].abs().max()).item())
    scale = np.abs(cell.params)
    scale[12:15] = 1.0
    dy_err = np.abs((a['dY'] - b['dY']) * scale).max() / np.abs(b['dY'] * scale).max()
    q_err = np.abs(a['Q'] - b['Q']).max() / np.abs(b['Q']).max()
    _say(precision, 'lambda', lam_err, 'dY', dy_err, 'Q', q_err, 'steps', a['ir_steps'], b['ir_steps'], b['adj_steps'])
    assert lam_err <= 1e-8 and dy_err <= 1e-9 and q_err <= 1e-12
