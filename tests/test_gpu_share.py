"""Factorization sharing (``FrequencySolver.solve_batch(share=k)``, ``sawfem_mf_solve_shared``):
the identity slot map reproduces the plain solve bit for bit, and shared factors refined against
each frequency's own FP64 operator meet the same accuracy as one factorization per frequency."""
import sys

import numpy as np
import pytest

from sawsim_amd import _lib, workloads
from sawsim_amd.piezo_fem import stream_ptr

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_identity_slot_map_matches_the_plain_solve():
    _say('[test_identity_slot_map_matches_the_plain_solve]')
    cell = workloads.config3(mesh_um=0.12, points=6)
    eng = workloads.build_engine(cell)
    torch = eng.torch
    om = torch.as_tensor(2 * np.pi * cell.frequencies(), device=eng.dev)
    B, n = om.numel(), eng.n_active
    arena = eng._arena(eng.mf.arena_elems * B, torch.complex64)
    uarena = torch.empty(eng.mf.uarena_elems * B * 2, dtype=torch.complex128, device=eng.dev)
    s = stream_ptr()
    _lib.call('sawfem_mf_factor', eng._plan, s, 0, B, om.data_ptr(), eng.beta, eng.Kaa.data_ptr(), eng.Maa.data_ptr(),
              eng._dptr(eng.Daa), arena.data_ptr(), 64)
    x0 = torch.randn((B, 2, n), dtype=torch.complex128, device=eng.dev,
                     generator=torch.Generator(device=eng.dev).manual_seed(1))
    a, b = x0.clone(), x0.clone()
    ident = torch.arange(B, dtype=torch.int32, device=eng.dev)
    for tr in (0, 1):
        _lib.call('sawfem_mf_solve', eng._plan, s, 0, B, 2, arena.data_ptr(), uarena.data_ptr(), a.data_ptr(), tr, 64,
                  None, 0)
        _lib.call('sawfem_mf_solve_shared', eng._plan, s, 0, B, 2, arena.data_ptr(), B, ident.data_ptr(),
                  uarena.data_ptr(), b.data_ptr(), tr, 64, None, 0)
        assert torch.equal(a, b)
    # an invalid slot count is refused before any launch
    rc = _lib.raw('sawfem_mf_solve_shared', eng._plan, s, 0, B, 2, arena.data_ptr(), B - 1, None, uarena.data_ptr(),
                  b.data_ptr(), 0, 64, None, 0)
    assert rc == _lib.INVALID


def _rows_solve(eng, s, B, R, arena, G, uarena, x0, sel, nsel, tail):
    """sawfem_mf_solve_groups_rows on a row-major copy of x0 ([B][R][n]), returned as [B][R][n]."""
    torch = eng.torch
    n = eng.n_active
    ld = -(-B * R // 64) * 64 + 64           # padded beyond the vectors: padding must stay untouched
    xr = torch.full((n, ld), 7.0 + 3.0j, dtype=torch.complex128, device=eng.dev)
    xr[:, :B * R] = x0.reshape(B * R, n).t()
    _lib.call('sawfem_mf_solve_groups_rows', eng._plan, s, B, R, arena.data_ptr(), G, uarena.data_ptr(),
              xr.data_ptr(), ld, 64, sel, nsel, tail)
    assert (xr[:, B * R:] == 7.0 + 3.0j).all()
    return xr[:, :B * R].t().reshape(B, R, n)


@pytest.mark.parametrize('G', [2, 3, 4, 8])
def test_grouped_solve_matches_the_slot_map_solve(G):
    """sawfem_mf_solve_groups (the vectors of a group of frequencies in one workgroup) against
    sawfem_mf_solve_shared with the slot map b -> b // G, for one and two vectors per frequency,
    both transposes, a partial last group and a selection of groups (unselected vectors
    untouched): bit for bit through the pivot-triangle and panel kernels at every block size
    (their per-vector operations do not depend on the block size); with the LDS block kernel
    on the one-workgroup fronts (blocks of 4, 8, 16 vectors; other summation order) within
    1e-10 of the largest entry of each vector.  Every call is repeated on the vectors stored
    row-major (sawfem_mf_solve_groups_rows): bit for bit the same."""
    _say(f'[test_grouped_solve_matches_the_slot_map_solve] G {G}')
    cell = workloads.config3(mesh_um=0.12, points=6)
    eng = workloads.build_engine(cell)
    torch = eng.torch
    B = 11
    f = np.linspace(cell.frequencies()[0], cell.frequencies()[-1], B)
    Bf = -(-B // G)
    grp = np.arange(B) // G
    om_f = torch.as_tensor(2 * np.pi * np.bincount(grp, weights=f) / np.bincount(grp), device=eng.dev)
    n = eng.n_active
    arena = eng._arena(eng.mf.arena_elems * Bf, torch.complex64)
    uarena = torch.empty(eng.mf.uarena_elems * B * 2, dtype=torch.complex128, device=eng.dev)
    s = stream_ptr()
    _lib.call('sawfem_mf_factor', eng._plan, s, 0, Bf, om_f.data_ptr(), eng.beta, eng.Kaa.data_ptr(),
              eng.Maa.data_ptr(), eng._dptr(eng.Daa), arena.data_ptr(), 64)
    slot = torch.as_tensor(grp.astype(np.int32), device=eng.dev)
    gen = torch.Generator(device=eng.dev).manual_seed(G)
    nfull = B // G
    for R in (1, 2):
        x0 = torch.randn((B, R, n), dtype=torch.complex128, device=eng.dev, generator=gen)
        for tr in (0, 1):
            ref = x0.clone()
            _lib.call('sawfem_mf_solve_shared', eng._plan, s, 0, B, R, arena.data_ptr(), Bf, slot.data_ptr(),
                      uarena.data_ptr(), ref.data_ptr(), tr, 64, None, 0)
            for rc in (1, 2, 4, 16):
                assert _lib.call_value('sawfem_mf_set_group_tri', eng._plan, rc, 0) == rc
                got = x0.clone()
                _lib.call('sawfem_mf_solve_groups', eng._plan, s, B, R, arena.data_ptr(), G, uarena.data_ptr(),
                          got.data_ptr(), tr, 64, None, 0, 1)
                assert torch.equal(got, ref), (R, tr, rc)
                if tr == 0:
                    assert torch.equal(_rows_solve(eng, s, B, R, arena, G, uarena, x0, None, 0, 1), got), (R, rc)
            _lib.call_value('sawfem_mf_set_group_tri', eng._plan, 4, 160 * 1024)
            for rc in (4, 8, 16):
                assert _lib.call_value('sawfem_mf_set_group_rc', eng._plan, rc) == rc
                got = x0.clone()
                _lib.call('sawfem_mf_solve_groups', eng._plan, s, B, R, arena.data_ptr(), G, uarena.data_ptr(),
                          got.data_ptr(), tr, 64, None, 0, 1)
                err = ((got - ref).abs().amax(dim=2) / ref.abs().amax(dim=2)).max().item()
                assert err <= 1e-10, (R, tr, rc, err)
                if tr == 0:
                    assert torch.equal(_rows_solve(eng, s, B, R, arena, G, uarena, x0, None, 0, 1), got), (R, rc)
            _lib.call_value('sawfem_mf_set_group_rc', eng._plan, 16)
            # groups 1 .. nfull-1 selected, the partial group not: group 0 and the tail untouched
            _lib.call_value('sawfem_mf_set_group_tri', eng._plan, 4, 0)
            sel = torch.arange(1, nfull, dtype=torch.int32, device=eng.dev)
            got = x0.clone()
            _lib.call('sawfem_mf_solve_groups', eng._plan, s, B, R, arena.data_ptr(), G, uarena.data_ptr(),
                      got.data_ptr(), tr, 64, sel.data_ptr() if sel.numel() else None, int(sel.numel()), 0)
            if sel.numel() == 0:      # a NULL list means every full group
                assert torch.equal(got[:nfull * G], ref[:nfull * G])
            else:
                assert torch.equal(got[:G], x0[:G])
                assert torch.equal(got[G:nfull * G], ref[G:nfull * G])
            assert torch.equal(got[nfull * G:], x0[nfull * G:])
            if tr == 0:
                gr = _rows_solve(eng, s, B, R, arena, G, uarena, x0, sel.data_ptr() if sel.numel() else None,
                                 int(sel.numel()), 0)
                assert torch.equal(gr, got)
    _lib.call_value('sawfem_mf_set_group_tri', eng._plan, 4, 48 * 1024)
    # invalid group sizes and selections are refused before any launch
    assert _lib.raw('sawfem_mf_solve_groups', eng._plan, s, B, 1, arena.data_ptr(), 0, uarena.data_ptr(),
                    x0.data_ptr(), 0, 64, None, 0, 1) == _lib.INVALID
    assert _lib.raw('sawfem_mf_solve_groups', eng._plan, s, B, 1, arena.data_ptr(), G, uarena.data_ptr(),
                    x0.data_ptr(), 0, 64, slot.data_ptr(), nfull + 1, 1) == _lib.INVALID
    # a row stride below the vector count is refused
    assert _lib.raw('sawfem_mf_solve_groups_rows', eng._plan, s, B, 1, arena.data_ptr(), G, uarena.data_ptr(),
                    x0.data_ptr(), B - 1, 64, None, 0, 1) == _lib.INVALID


@pytest.mark.parametrize('share', [2, 4])
def test_shared_factors_meet_the_fp64_accuracy(share):
    """Config 1 as the benchmark names it (51,752 unknowns), 24 consecutive frequencies of its
    band: with one factorization per `share` frequencies, mixed Y11 within 1e-6 and dY11/dp
    within 1e-5 of the FP64 path (refinement to 1e-8 as without sharing), fewer
    factorizations than frequencies."""
    _say(f'[test_shared_factors_meet_the_fp64_accuracy] share {share}')
    cell = workloads.config1()
    eng = workloads.build_engine(cell, gradient=True)
    f = cell.frequencies()[40:64]
    Q64, dY64, _ = eng.sweep_charges(f, precision='fp64', gradient=True)
    eng.set_share(share)
    Q, dY, info = eng.sweep_charges(f, precision='mixed', gradient=True, batch=f.size)
    scale = np.abs(cell.params)
    scale[12:15] = 1.0
    ey = np.abs(Q - Q64) / np.abs(Q64)
    eg = np.abs((dY - dY64) * scale).max(axis=1) / np.abs(dY64 * scale).max(axis=1)
    _say('Y11', ey.max(), 'dY', eg.max(), 'factorizations', info[0]['factorizations'], 'own', info[0]['own_factors'],
         'passes', info[0]['solved_per_pass'])
    assert info[0]['share'] == share and info[0]['factorizations'] < f.size
    assert ey.max() <= 1e-6 and eg.max() <= 1e-5
