"""GPU parity: HIP assembly, Bloch reduction, multifrontal LU and Y11 against the oracle."""
import sys

import numpy as np
import pytest

import helpers
import mf_emulator
from sawsim_amd import workloads
from sawsim_amd.piezo_fem import assemble, assemble_ME0, assemble_device
from sawsim_amd.sweep import FrequencySolver, bc_periodic

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_assemble(cell, omega_top, damping=False):
    fn = assemble if cell.ndisp == 3 else assemble_ME0
    if not damping:
        return fn(cell.model, cell.materials, cell.mat_idx, cell.pml, omega_top)
    nd = cell.ndisp + 1
    from sawsim_amd.materials import lift_projected
    mats = cell.materials if nd == 4 else [lift_projected(m) for m in cell.materials]
    mask = np.zeros(len(mats), dtype=bool)
    mask[0] = True
    return assemble_device(cell.model, mats, cell.mat_idx, cell.pml, omega_top, nd, damping_mask=mask)


def _engine(cell, omega_top, **kw):
    K, M, D = _gpu_assemble(cell, omega_top, damping=True)
    bcs = bc_periodic(cell.model, cell.mesh.lines, ndisp=cell.ndisp)
    eng = FrequencySolver(K, M, *bcs, **kw)
    if cell.beta_dk:
        eng.set_damping(D, cell.beta_dk)
    return eng


CELLS = {
    'double_layer': lambda: workloads.double_layer(),
    'config1_coarse': lambda: workloads.config1(mesh_um=0.09, points=5),
    'config3_coarse': lambda: workloads.config3(mesh_um=0.09, points=5),
}


@pytest.mark.parametrize('name', list(CELLS))
def test_assembly_matches_oracle(name):
    _say(f'[test_assembly_matches_oracle] {name}')
    cell = CELLS[name]()
    wtop = 2 * np.pi * cell.band_hz[1]
    K, M, D = _gpu_assemble(cell, wtop, damping=True)
    Ko, Mo = helpers.oracle_assemble(cell, wtop)
    Do, _ = helpers.oracle_assemble(cell, wtop, helpers.oracle_damping_materials(cell))
    for gpu, ref in ((K, Ko), (M, Mo), (D, Do)):
        g = gpu.to_scipy()
        assert g.shape == ref.shape
        # oracle pattern is contained in the structural pattern; extra slots are exact zeros
        diff = (g - ref).tocsr()
        scale = abs(ref).max()
        assert abs(diff).max() <= 1e-12 * scale
        gz = g.copy(); gz.data[:] = 1
        rz = ref.copy(); rz.data[:] = 1
        assert (rz - rz.multiply(gz)).nnz == 0


@pytest.mark.parametrize('name', list(CELLS))
def test_bloch_maps_and_active_matrix(name):
    _say(f'[test_bloch_maps_and_active_matrix] {name}')
    cell = CELLS[name]()
    wtop = 2 * np.pi * cell.band_hz[1]
    eng = _engine(cell, wtop)
    ref = helpers.oracle_engine(cell, wtop)
    assert np.array_equal(eng.active, ref.active)
    o = np.argsort(ref.slaves)
    assert np.array_equal(eng.slaves, ref.slaves[o]) and np.array_equal(eng.masters, ref.masters[o])
    r = np.repeat(np.arange(eng.n_active), np.diff(eng.bloch.aa_indptr))
    c = eng.bloch.aa_indices
    Kaa = eng.Kaa.cpu().numpy()
    Kref = np.asarray(ref.Km_aa[r, c]).ravel()
    assert np.abs(Kaa - Kref).max() <= 1e-12 * np.abs(Kref).max()
    # every oracle nonzero is a slot of ours
    assert ref.Km_aa.nnz <= eng.bloch.nnz_aa
    kts = np.concatenate([eng.kts_a.cpu().numpy(), [eng.kts_z]])
    disp = np.zeros(ref.gdof, dtype=complex)
    x = np.random.default_rng(0).standard_normal(eng.n_active) + 0j
    Qo = ref.kts @ ref.expand(x)
    Qg = kts[:-1] @ x + kts[-1]
    assert abs(Qg - Qo) <= 1e-12 * abs(Qo) + 1e-30


@pytest.mark.parametrize('path', ['fused', 'panel'])
@pytest.mark.parametrize('prec,nb', [('fp64', 32), ('mixed', 32), ('mixed', 64)])
def test_factor_arena_matches_emulation(prec, nb, path):
    """Every front of the GPU factorization against the NumPy emulation of the kernels'
    data flow; 'fused': fronts up to the LDS limit go through the fused small-front kernel,
    'panel': every front through the panel kernels."""
    _say(f'[test_factor_arena_matches_emulation] {prec} nb={nb} {path}')
    import torch
    cell = workloads.double_layer()
    wtop = 2 * np.pi * cell.band_hz[1]
    eng = _engine(cell, wtop, leaf_weight=160)   # larger leaves: several panels per front
    omegas = 2 * np.pi * np.array([1.7e9, 1.85e9, 2.2e9])
    B = omegas.size
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    pr = 0 if prec == 'mixed' else 1
    lim = _lib.call_value('sawfem_mf_set_small', eng._plan, pr, 128 if path == 'fused' else 0)
    nsmall = int((eng.mf.m <= lim).sum())
    _say(f'fused threshold {lim}: {nsmall} of {eng.mf.nf} fronts')
    assert (nsmall > 0 and nsmall < eng.mf.nf) if path == 'fused' else nsmall == 0
    arena = torch.empty(eng.mf.arena_elems * B, dtype=torch.complex64 if pr == 0 else torch.complex128, device='cuda')
    om = torch.as_tensor(omegas, device='cuda')
    _lib.call('sawfem_mf_factor', eng._plan, stream_ptr(), pr, B, om.data_ptr(), 0.0, eng.Kaa.data_ptr(),
              eng.Maa.data_ptr(), None, arena.data_ptr(), nb)
    got = arena.cpu().numpy().astype(np.complex128)
    ref = mf_emulator.factor(eng.mf, B, omegas, 0.0, eng.Kaa.cpu().numpy(), eng.Maa.cpu().numpy(), None, nb=nb)
    off = eng.mf.off
    # intermediate fronts without pivoting: roundoff differences (operation order) are amplified by
    # pivot growth in a few fronts; indexing mistakes give O(1) errors
    tol = 1e-6 if pr == 1 else 5e-3
    for f in range(eng.mf.nf):
        m = int(eng.mf.m[f])
        a = got[off[f] * B: off[f] * B + B * m * m]
        b = ref[off[f] * B: off[f] * B + B * m * m]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert err <= tol, (f, m, int(eng.mf.p[f]), err)


@pytest.mark.parametrize('name', list(CELLS))
def test_y11_fp64_and_mixed_vs_oracle(name):
    _say(f'[test_y11_fp64_and_mixed_vs_oracle] {name}')
    cell = CELLS[name]()
    wtop = 2 * np.pi * cell.band_hz[1]
    eng = _engine(cell, wtop)
    ref = helpers.oracle_engine(cell, wtop)
    f = cell.frequencies(5)
    Q64, info64 = eng.sweep_charges(f, precision='fp64')
    Qmx, infomx = eng.sweep_charges(f, precision='mixed')
    Qo = np.array([helpers.refined_charge(ref, 2 * np.pi * fz)[0] for fz in f])
    e64 = np.abs(Q64 - Qo) / np.abs(Qo)
    emx = np.abs(Qmx - Q64) / np.abs(Q64)
    _say(name, 'fp64 vs oracle', e64.max(), 'mixed vs fp64', emx.max(), info64, infomx)
    assert e64.max() <= 1e-8
    assert emx.max() <= 1e-6


def test_double_layer_reproduces_stored_curve():
    """|Y| of the reference's stored sp_double_layer curve at its test frequencies."""
    _say('[test_double_layer_reproduces_stored_curve]')
    import json
    from pathlib import Path
    data = Path(__file__).parent / 'data'
    meta = json.loads((data / 'sp_double_layer.json').read_text())
    d = np.load(data / 'sp_double_layer.npz')
    idx = np.array(meta['test_indices'])
    f = d['frequency_hz'][idx]
    cell = workloads.double_layer(points=len(f), band=(f[0], f[-1]))
    eng = _engine(cell, 2 * np.pi * f[-1])
    for prec in ('fp64', 'mixed'):
        Q, _ = eng.sweep_charges(f, precision=prec)
        y = -1j * 2 * np.pi * f * Q
        err = np.abs(np.abs(y) - d['sawsim_magnitude'][idx]) / d['sawsim_magnitude'][idx]
        _say(prec, err)
        assert err.max() <= meta['regression_rtol']
        assert err.max() <= 1e-6


def test_gpu_assembly_matches_reference_golden_tiny_me0():
    """GPU K, M, D against the reference's own assemble_ME0 output (golden vectors)."""
    _say('[test_gpu_assembly_matches_reference_golden_tiny_me0]')
    import scipy.sparse as sp
    from golden_cells import tiny_cell
    cell, g, p = tiny_cell()
    n = 3 * cell.model.N
    K, M, D = _gpu_assemble(cell, float(g['omega_top']), damping=True)
    for dev, name in ((K, 'K'), (M, 'M'), (D, 'D')):
        ref = sp.csr_matrix((g[name + '_data'], g[name + '_indices'], g[name + '_indptr']), shape=(n, n))
        mine = dev.to_scipy()
        if name == 'K':    # K keeps every element slot in the reference too: identical pattern
            assert np.array_equal(mine.indptr, ref.indptr) and np.array_equal(mine.indices, ref.indices)
        # M and D: the reference pattern is contained in ours, extra slots hold exact zeros
        assert abs(mine - ref).max() <= 1e-12 * abs(ref).max()
        rz = ref.copy(); rz.data[:] = 1
        mz = mine.copy(); mz.data[:] = 1
        assert (rz - rz.multiply(mz)).nnz == 0
    eng = FrequencySolver(K, M, *bc_periodic(cell.model, cell.mesh.lines, ndisp=2))
    assert np.array_equal(eng.active, g['active'])


def test_solve_on_a_frequency_subset():
    """sawfem_mf_solve with a frequency list: the listed frequencies get the full solve's
    result bit for bit, the others are left untouched."""
    _say('[test_solve_on_a_frequency_subset]')
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    cell = workloads.double_layer()
    eng = _engine(cell, 2 * np.pi * cell.band_hz[1])
    omegas = 2 * np.pi * np.linspace(1.6e9, 2.4e9, 6)
    B, n = omegas.size, eng.n_active
    om = torch.as_tensor(omegas, device='cuda')
    arena = torch.empty(eng.mf.arena_elems * B, dtype=torch.complex64, device='cuda')
    uarena = torch.empty(eng.mf.uarena_elems * B * 2, dtype=torch.complex128, device='cuda')
    s = stream_ptr()
    _lib.call('sawfem_mf_factor', eng._plan, s, 0, B, om.data_ptr(), 0.0, eng.Kaa.data_ptr(), eng.Maa.data_ptr(),
              None, arena.data_ptr(), 64)
    g = torch.Generator(device='cpu').manual_seed(3)
    rhs = torch.complex(torch.randn(B, 2, n, generator=g, dtype=torch.float64),
                        torch.randn(B, 2, n, generator=g, dtype=torch.float64)).cuda()
    full = rhs.clone()
    _lib.call('sawfem_mf_solve', eng._plan, s, 0, B, 2, arena.data_ptr(), uarena.data_ptr(), full.data_ptr(), 0, 64,
              None, 0)
    part = rhs.clone()
    sel = torch.tensor([1, 4], dtype=torch.int32, device='cuda')
    _lib.call('sawfem_mf_solve', eng._plan, s, 0, B, 2, arena.data_ptr(), uarena.data_ptr(), part.data_ptr(), 0, 64,
              sel.data_ptr(), 2)
    torch.cuda.synchronize()
    for b in range(B):
        want = full[b] if b in (1, 4) else rhs[b]
        assert torch.equal(part[b], want), b


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_one_workgroup_solves_match_the_kernel_pair(precision):
    """Solving the small fronts with one workgroup each (pivot triangle + panel together) gives
    the same sweep as the pivot-triangle + panel kernel pair, forward and paired adjoint."""
    _say(f'[test_one_workgroup_solves_match_the_kernel_pair] {precision}')
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell, gradient=True)
    om = 2 * np.pi * cell.frequencies(4)
    assert eng.set_solve_full(0) == 0
    a = eng.solve_batch(om, precision=precision, gradient=True)
    assert eng.set_solve_full(400) == 400
    b = eng.solve_batch(om, precision=precision, gradient=True)
    eng.set_solve_full(0)
    tol = 1e-12 if precision == 'fp64' else 1e-9
    assert np.max(np.abs(a['Q'] - b['Q']) / np.abs(a['Q'])) <= tol
    assert np.max(np.abs(a['dY'] - b['dY']).max(axis=1) / np.abs(a['dY']).max(axis=1)) <= 100 * tol


def test_factor_and_solve_on_a_side_torch_stream():
    """The plan forks the size buckets of a level onto its own streams and joins them back to
    the caller's stream: factor + solve issued on a non-default torch stream give the same
    arena and solution bit for bit as on the default stream."""
    _say('[test_factor_and_solve_on_a_side_torch_stream]')
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    cell = workloads.double_layer()
    eng = _engine(cell, 2 * np.pi * cell.band_hz[1])
    # some level has fronts in more than one size bucket (m <= 64, 128, 192, ...): the fork runs
    edges = [64, 128, 192, 320, 576, 1088]
    mf = eng.mf
    buckets = [np.unique(np.searchsorted(edges, mf.m[mf.lvl_fronts[mf.lvl_ptr[L]:mf.lvl_ptr[L + 1]]]))
               for L in range(mf.lvl_ptr.size - 1)]
    assert max(b.size for b in buckets) > 1
    omegas = 2 * np.pi * np.linspace(1.6e9, 2.4e9, 5)
    B, n = omegas.size, eng.n_active
    om = torch.as_tensor(omegas, device='cuda')
    g = torch.Generator(device='cpu').manual_seed(5)
    rhs = torch.complex(torch.randn(B, 2, n, generator=g, dtype=torch.float64),
                        torch.randn(B, 2, n, generator=g, dtype=torch.float64)).cuda()

    def run():
        s = stream_ptr()
        arena = torch.empty(eng.mf.arena_elems * B, dtype=torch.complex64, device='cuda')
        uarena = torch.empty(eng.mf.uarena_elems * B * 2, dtype=torch.complex128, device='cuda')
        x = rhs.clone()
        _lib.call('sawfem_mf_factor', eng._plan, s, 0, B, om.data_ptr(), 0.0, eng.Kaa.data_ptr(),
                  eng.Maa.data_ptr(), None, arena.data_ptr(), 64)
        _lib.call('sawfem_mf_solve', eng._plan, s, 0, B, 2, arena.data_ptr(), uarena.data_ptr(), x.data_ptr(), 0,
                  64, None, 0)
        return arena, x

    torch.cuda.synchronize()
    a0, x0 = run()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        a1, x1 = run()
    torch.cuda.synchronize()
    assert torch.equal(a0, a1)
    assert torch.equal(x0, x1)


def test_frequency_major_and_transposed_residuals_agree():
    """sawfem_residual ([B][n] vectors, eight frequencies per thread) and sawfem_residual_t
    ([n][Bp] vectors, one wave per row and 64 frequencies) compute the same FP64 residual, and
    both match r = rhs - A(w) x assembled on the host from the same values."""
    _say('[test_frequency_major_and_transposed_residuals_agree]')
    import scipy.sparse as sp
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell)
    n, B = eng.n_active, 11                       # 11 = one full and one partial group of 8
    Bp = 64
    omegas = 2 * np.pi * np.linspace(1.35e9, 1.5e9, B)
    om = torch.as_tensor(omegas, device='cuda')
    g = torch.Generator(device='cpu').manual_seed(7)
    x = torch.complex(torch.randn(B, n, generator=g, dtype=torch.float64),
                      torch.randn(B, n, generator=g, dtype=torch.float64)).cuda()
    rhs = torch.complex(torch.randn(B, n, generator=g, dtype=torch.float64),
                        torch.randn(B, n, generator=g, dtype=torch.float64)).cuda()
    s = stream_ptr()
    r1 = torch.empty_like(x)
    _lib.call('sawfem_residual', s, n, B, eng.aa_indptr.data_ptr(), eng.aa_indices.data_ptr(), None,
              eng.Kaa.data_ptr(), eng.Maa.data_ptr(), eng._dptr(eng.Daa), om.data_ptr(), eng.beta, x.data_ptr(),
              rhs.data_ptr(), r1.data_ptr())
    xt = torch.zeros((n, Bp), dtype=torch.complex128, device='cuda')
    rt = torch.zeros((n, Bp), dtype=torch.complex128, device='cuda')
    rhst = torch.zeros((n, Bp), dtype=torch.complex128, device='cuda')
    xt[:, :B] = x.T
    rhst[:, :B] = rhs.T
    _lib.call('sawfem_residual_t', s, n, B, Bp, eng.aa_indptr.data_ptr(), eng.aa_indices.data_ptr(), None,
              eng.Kaa.data_ptr(), eng.Maa.data_ptr(), eng._dptr(eng.Daa), om.data_ptr(), eng.beta, xt.data_ptr(),
              rhst.data_ptr(), rt.data_ptr())
    torch.cuda.synchronize()
    r2 = rt[:, :B].T
    scale = r1.abs().max().item()
    assert (r1 - r2).abs().max().item() <= 1e-12 * scale
    ip, ix = eng.aa_indptr.cpu().numpy(), eng.aa_indices.cpu().numpy()
    K = sp.csr_matrix((eng.Kaa.cpu().numpy(), ix, ip), shape=(n, n))
    M = sp.csr_matrix((eng.Maa.cpu().numpy(), ix, ip), shape=(n, n))
    D = sp.csr_matrix((eng.Daa.cpu().numpy(), ix, ip), shape=(n, n)) if eng.Daa is not None else 0 * K
    xs, rs = x.cpu().numpy(), rhs.cpu().numpy()
    for b in (0, 7, 8, 10):
        w = omegas[b]
        ref = rs[b] - (K + 1j * w * eng.beta * D - w * w * M) @ xs[b]
        assert np.abs(r1[b].cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()


def test_edge_cases_empty_single_and_voltage_scaling():
    """No frequencies: empty results and no launch; one frequency: the same charge as inside a
    batch; Q is linear in the drive voltage (V = 2 doubles every right-hand side exactly, so the
    charges double to rounding); an unknown precision is refused."""
    _say('[test_edge_cases_empty_single_and_voltage_scaling]')
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell, gradient=True)
    Q0, dY0, info0 = eng.sweep_charges([], precision='mixed', gradient=True)
    assert Q0.shape == (0,) and dY0.shape == (0, eng.n_params) and info0 == []
    Qe, infoe = eng.sweep_charges(np.zeros(0), precision='fp64')
    assert Qe.shape == (0,) and infoe == []
    f = np.linspace(1.35e9, 1.5e9, 5)
    Q1, dY1, _ = eng.sweep_charges(f, precision='mixed', gradient=True)
    Qs, dYs, _ = eng.sweep_charges(f[2:3], precision='mixed', gradient=True)
    _say('single vs batch', abs(Qs[0] - Q1[2]) / abs(Q1[2]))
    assert abs(Qs[0] - Q1[2]) <= 1e-7 * abs(Q1[2])
    assert np.abs(dYs[0] - dY1[2]).max() <= 1e-6 * np.abs(dY1[2]).max()
    Qv1, _ = eng.sweep_charges(f, V=1.0, precision='mixed')
    Qv2, _ = eng.sweep_charges(f, V=2.0, precision='mixed')
    _say('V=2 vs 2 x V=1', np.max(np.abs(Qv2 - 2 * Qv1) / np.abs(2 * Qv1)))
    assert np.allclose(Qv2, 2 * Qv1, rtol=1e-12, atol=0)
    with pytest.raises(ValueError):
        eng.sweep_charges(f, precision='fp32')
    with pytest.raises(ValueError):
        eng.solve_batch([], precision='mixed')


@pytest.mark.parametrize('transposed,B', [(False, 40), (False, 70), (True, 70), (False, 520), (True, 520)])
def test_grouped_residual_matches_the_row_residual(transposed, B):
    """sawfem_residual_sv (one wave per merged node's rows, each x row read once for the group,
    1, 2 or 8 frequency tiles per wave) equals sawfem_residual_t bit for bit, for A and, with the
    transposed-slot map, A^T; on the config-3 coarse cell (groups of three active unknowns), with
    1, 2 and 9 tiles of 64 frequencies (the last column of 8 tiles holding one) and the padded
    frequencies of the last tile."""
    _say(f'[test_grouped_residual_matches_the_row_residual] transposed={transposed} B={B}')
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell, gradient=transposed)
    n, Bp = eng.n_active, -(-B // 64) * 64
    gptr, grows, ng = eng.res_groups
    assert ng < n
    om = torch.as_tensor(2 * np.pi * np.linspace(1.35e9, 1.5e9, B), device='cuda')
    g = torch.Generator(device='cpu').manual_seed(11)
    xt = torch.complex(torch.randn(n, Bp, generator=g, dtype=torch.float64),
                       torch.randn(n, Bp, generator=g, dtype=torch.float64)).cuda()
    rhst = torch.complex(torch.randn(n, Bp, generator=g, dtype=torch.float64),
                         torch.randn(n, Bp, generator=g, dtype=torch.float64)).cuda()
    ts = eng._dptr(eng.tslot) if transposed else None
    s = stream_ptr()
    r1 = torch.zeros_like(xt)
    r2 = torch.zeros_like(xt)
    _lib.call('sawfem_residual_t', s, n, B, Bp, eng.aa_indptr.data_ptr(), eng.aa_indices.data_ptr(), ts,
              eng.Kaa.data_ptr(), eng.Maa.data_ptr(), eng._dptr(eng.Daa), om.data_ptr(), eng.beta, xt.data_ptr(),
              rhst.data_ptr(), r1.data_ptr())
    _lib.call('sawfem_residual_sv', s, ng, B, Bp, None, gptr.data_ptr(), grows.data_ptr(), eng.aa_indptr.data_ptr(),
              eng.aa_indices.data_ptr(), ts, eng.Kaa.data_ptr(), eng.Maa.data_ptr(), eng._dptr(eng.Daa),
              om.data_ptr(), eng.beta, xt.data_ptr(), rhst.data_ptr(), r2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(r1[:, :B], r2[:, :B])
    assert not r2[:, B:].any()        # the padded columns of the last tile are not touched
    # a column map: a third of the columns in shuffled order, each equal to the full residual's
    # column, every other column untouched
    cols = np.random.default_rng(B).permutation(B)[:max(1, B // 3)].astype(np.int32)
    cmap = torch.as_tensor(cols, device='cuda')
    r3 = torch.zeros_like(xt)
    _lib.call('sawfem_residual_sv', s, ng, cols.size, Bp, cmap.data_ptr(), gptr.data_ptr(), grows.data_ptr(),
              eng.aa_indptr.data_ptr(), eng.aa_indices.data_ptr(), ts, eng.Kaa.data_ptr(), eng.Maa.data_ptr(),
              eng._dptr(eng.Daa), om.data_ptr(), eng.beta, xt.data_ptr(), rhst.data_ptr(), r3.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(r3[:, cmap.long()], r1[:, cmap.long()])
    rest = np.setdiff1d(np.arange(Bp), cols)
    assert not r3[:, torch.as_tensor(rest, device='cuda')].any()


def test_weighted_dot_matches_torch():
    """sawfem_wdot (the refinement's duality products in one pass) against torch's product and sum,
    strided rows of a [B][2][n] array, with and without a selection and weights."""
    _say('[test_weighted_dot_matches_torch]')
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    g = torch.Generator(device='cpu').manual_seed(5)
    B, n = 9, 50_003
    X = torch.complex(torch.randn(B, 2, n, generator=g, dtype=torch.float64),
                      torch.randn(B, 2, n, generator=g, dtype=torch.float64)).cuda()
    c = torch.complex(torch.randn(B, n, generator=g, dtype=torch.float64),
                      torch.randn(B, n, generator=g, dtype=torch.float64)).cuda()
    w = torch.rand(n, generator=g, dtype=torch.float64).cuda() + 0.5
    out = torch.empty(B, dtype=torch.complex128, device='cuda')
    s = stream_ptr()
    _lib.call('sawfem_wdot', s, n, B, X.data_ptr() + 16 * n, 2 * n, w.data_ptr(), c.data_ptr(), n, None, out.data_ptr())
    ref = (X[:, 1] * w * c).sum(dim=1)
    assert ((out - ref).abs().max() / ref.abs().max()).item() <= 1e-13
    sel = torch.tensor([7, 2, 4], dtype=torch.int32, device='cuda')
    _lib.call('sawfem_wdot', s, n, 3, X.data_ptr(), 2 * n, None, X.data_ptr() + 16 * n, 2 * n, sel.data_ptr(),
              out.data_ptr())
    ref = (X[:, 0] * X[:, 1]).sum(dim=1)[sel.long()]
    assert ((out[:3] - ref).abs().max() / ref.abs().max()).item() <= 1e-13


def test_row_major_charge_and_weighted_dot_match_bit_for_bit():
    """sawfem_charge_rows / sawfem_wdot_rows / sawfem_wdot_pairs_rows on the vectors of a [B][2][n]
    array stored row-major ([n][ld], column k = vector k) equal sawfem_charge / sawfem_wdot on the
    array itself bit for bit (the same summation order), with the identity and with shuffled
    column maps, weights and no weights."""
    _say('[test_row_major_charge_and_weighted_dot_match_bit_for_bit]')
    import torch
    from sawsim_amd import _lib
    from sawsim_amd.piezo_fem import stream_ptr
    g = torch.Generator(device='cpu').manual_seed(8)
    B, n = 37, 70_001
    nv, ld = 2 * B, 128
    cx = lambda *shape: torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                                      torch.randn(*shape, generator=g, dtype=torch.float64)).cuda()
    X, C, kts = cx(B, 2, n), cx(B, 2, n), cx(n)
    w = torch.rand(n, generator=g, dtype=torch.float64).cuda() + 0.5
    Xr = torch.zeros(n, ld, dtype=torch.complex128, device='cuda')
    Xr[:, :nv] = X.reshape(nv, n).t()
    Cr = torch.zeros_like(Xr)
    Cr[:, :nv] = C.reshape(nv, n).t()
    work = torch.empty(1024 * nv, dtype=torch.complex128, device='cuda')
    s = stream_ptr()
    q1 = torch.empty(nv, dtype=torch.complex128, device='cuda')
    q2 = torch.empty_like(q1)
    _lib.call('sawfem_charge', s, n, nv, kts.data_ptr(), 0.25, -1.5, 2.0, X.data_ptr(), q1.data_ptr())
    _lib.call('sawfem_charge_rows', s, n, nv, ld, kts.data_ptr(), 0.25, -1.5, 2.0, Xr.data_ptr(), None,
              work.data_ptr(), q2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(q1, q2)
    cols = torch.tensor([9, 0, 73, 40], dtype=torch.int32, device='cuda')
    _lib.call('sawfem_charge_rows', s, n, 4, ld, kts.data_ptr(), 0.25, -1.5, 2.0, Xr.data_ptr(), cols.data_ptr(),
              work.data_ptr(), q2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(q1[cols.long()], q2[:4])
    even = torch.arange(0, nv, 2, dtype=torch.int32, device='cuda')
    odd = even + 1
    for wp in (w.data_ptr(), None):
        o1 = torch.empty(B, dtype=torch.complex128, device='cuda')
        o2 = torch.empty_like(o1)
        # a = vector 2b + 1 of X, c = vector 2b of C (strides of the [B][2][n] array)
        _lib.call('sawfem_wdot', s, n, B, X.data_ptr() + 16 * n, 2 * n, wp, C.data_ptr(), 2 * n, None, o1.data_ptr())
        _lib.call('sawfem_wdot_rows', s, n, B, ld, Xr.data_ptr(), odd.data_ptr(), wp, Cr.data_ptr(), even.data_ptr(),
                  work.data_ptr(), o2.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(o1, o2)
        # both products of a frequency's column pair in one pass (sawfem_wdot_pairs_rows)
        o3 = torch.empty_like(o1)
        p1, p2 = torch.empty_like(o1), torch.empty_like(o1)
        _lib.call('sawfem_wdot', s, n, B, X.data_ptr(), 2 * n, wp, C.data_ptr() + 16 * n, 2 * n, None, o3.data_ptr())
        _lib.call('sawfem_wdot_pairs_rows', s, n, B, ld, Xr.data_ptr(), wp, Cr.data_ptr(), work.data_ptr(),
                  p1.data_ptr(), p2.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(p1, o1) and torch.equal(p2, o3)
    # the refinement update with its measurements in one pass (sawfem_rows_update_measure) against
    # sawfem_axpy, sawfem_charge_rows and sawfem_wdot_rows one after another
    Rr = torch.zeros_like(Xr)
    Rr[:, :nv] = cx(nv, n).t()
    X1, X2 = Xr.clone(), Xr.clone()
    qa, qb = torch.empty(nv, dtype=torch.complex128, device='cuda'), torch.empty(nv, dtype=torch.complex128, device='cuda')
    ma, mb = torch.empty(B, dtype=torch.complex128, device='cuda'), torch.empty(B, dtype=torch.complex128, device='cuda')
    work3 = torch.empty(3072 * B, dtype=torch.complex128, device='cuda')
    _lib.call('sawfem_axpy', s, X1.numel(), X1.data_ptr(), Rr.data_ptr())
    _lib.call('sawfem_charge_rows', s, n, nv, ld, kts.data_ptr(), 0.25, -1.5, 2.0, X1.data_ptr(), None,
              work.data_ptr(), qa.data_ptr())
    _lib.call('sawfem_wdot_rows', s, n, B, ld, X1.data_ptr(), odd.data_ptr(), w.data_ptr(), Cr.data_ptr(),
              even.data_ptr(), work.data_ptr(), ma.data_ptr())
    _lib.call('sawfem_rows_update_measure', s, n, B, ld, X2.data_ptr(), Rr.data_ptr(), kts.data_ptr(), 0.25, -1.5, 2.0,
              w.data_ptr(), Cr.data_ptr(), work3.data_ptr(), qb.data_ptr(), mb.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(X1, X2) and torch.equal(qa, qb) and torch.equal(ma, mb)


@pytest.mark.parametrize('transposed', [False, True])
def test_compact_residual_matches_the_full_residual(transposed):
    """A refinement step with few active frequencies gathers their vectors (sawfem_transpose_map),
    evaluates rhs - A x for them only and scatters back: the active rows equal the full residual's
    bit for bit, the other rows are left untouched."""
    _say(f'[test_compact_residual_matches_the_full_residual] transposed={transposed}')
    import torch
    from sawsim_amd.sweep import _ResidualOp
    cell = workloads.config3(mesh_um=0.12, points=4)
    eng = workloads.build_engine(cell, gradient=transposed)
    n, Bf, per = eng.n_active, 37, 2
    om = torch.as_tensor(2 * np.pi * np.linspace(1.35e9, 1.5e9, Bf), device='cuda').repeat_interleave(per)
    g = torch.Generator(device='cpu').manual_seed(3)
    X = torch.complex(torch.randn(Bf, per, n, generator=g, dtype=torch.float64),
                      torch.randn(Bf, per, n, generator=g, dtype=torch.float64)).cuda()
    rhs = torch.complex(torch.randn(Bf * per, n, generator=g, dtype=torch.float64),
                        torch.randn(Bf * per, n, generator=g, dtype=torch.float64)).cuda()
    ts = eng.tslot if transposed else None
    res = _ResidualOp(eng, Bf * per, om)
    rhst = res.to_t(rhs)
    full = torch.empty_like(X)
    res(X, rhst, full, ts)
    active = np.zeros(Bf, dtype=bool)
    active[[0, 5, 6, 20, 36]] = True
    part = torch.full_like(X, 7.0)
    res(X, rhst, part, ts, active, per)
    torch.cuda.synchronize()
    a = torch.as_tensor(active, device='cuda')
    assert torch.equal(part[a], full[a])
    assert bool((part[~a] == 7.0).all())
