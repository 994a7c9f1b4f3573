"""GPU: sawfem_adjoint_elem_noalloc (element sensitivity sums in a static device array) against
sawfem_adjoint_elem (caller workspace) through the C ABI: two calls in a row, the second on a
smaller input, bit-identical outputs; a request above the array's size is refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from sawsim_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _args(eng, ne0, B, x, lam, W, work=None):
    from sawsim_amd.piezo_fem import SCALE, stream_ptr
    g = eng._grad
    tags, kinds, par = g['pml']
    ndu = eng.layout.nd - 1
    head = [stream_ptr(), ne0, g['elems'].data_ptr(), g['conn'].data_ptr(), g['xy'].data_ptr(), g['tag'].data_ptr(),
            g['npml'], tags.ctypes.data, kinds.ctypes.data, par.ctypes.data, g['omega_top'], ndu, g['N'],
            g['pos_rep'].data_ptr(), g['zmask'].data_ptr(), 1.0, SCALE, B, eng.n_active, x.data_ptr(),
            lam.data_ptr()]
    return head + ([work.data_ptr()] if work is not None else []) + [W.data_ptr()]


@pytest.mark.parametrize('ndisp_cell', ['me0', 'me1'])
def test_noalloc_adjoint_matches_the_workspace_entry_point(ndisp_cell):
    import torch
    _say(f'[test_noalloc_adjoint_matches_the_workspace_entry_point] {ndisp_cell}')
    if ndisp_cell == 'me0':
        cell = workloads.config3(mesh_um=0.12, points=3)
        eng = workloads.build_engine(cell, gradient=True)
    else:
        from sawsim_amd.materials import lift_projected  # noqa: F401  (ME1 cell: three displacements)
        cell = workloads.double_layer(points=3)
        eng = workloads.build_engine(cell)
        from sawsim_amd.params import from_constants
        eng.enable_gradient(cell.model, cell.mat_idx, cell.pml, 2 * np.pi * cell.band_hz[1],
                            from_constants(workloads.SP_BASELINE_3M, (0, -42, 0)), plane=None)
    ndu = eng.layout.nd - 1
    NT = (2 * ndu) ** 2 + 4 * ndu + 4
    n = eng.n_active
    gen = torch.Generator(device='cuda').manual_seed(0)
    ne_all = eng._grad['ne0']
    for B, ne0 in ((7, ne_all), (3, ne_all // 2)):           # second call: smaller input
        x = torch.randn((B, n), dtype=torch.complex128, device='cuda', generator=gen)
        lam = torch.randn((B, n), dtype=torch.complex128, device='cuda', generator=gen)
        W_ref = torch.empty(B * NT, dtype=torch.complex128, device='cuda')
        W_new = torch.full((B * NT,), complex(np.nan, np.nan), dtype=torch.complex128, device='cuda')
        work = torch.empty(B * ne0 * NT, dtype=torch.complex128, device='cuda')
        _lib.call('sawfem_adjoint_elem', *_args(eng, ne0, B, x, lam, W_ref, work))
        _lib.call('sawfem_adjoint_elem_noalloc', *_args(eng, ne0, B, x, lam, W_new))
        torch.cuda.synchronize()
        a, b = W_ref.cpu().numpy(), W_new.cpu().numpy()
        _say(f'B={B} ne0={ne0}: max |diff| {np.abs(a - b).max():.3e}, max |W| {np.abs(a).max():.3e}')
        assert np.all(np.isfinite(a)) and np.array_equal(a, b)
    # above the 64 MiB array: refused on the host before any launch
    B = (64 << 20) // (16 * NT * ne_all) + 1
    x = torch.zeros((B, n), dtype=torch.complex128, device='cuda')
    W = torch.empty(B * NT, dtype=torch.complex128, device='cuda')
    rc = _lib.raw('sawfem_adjoint_elem_noalloc', *_args(eng, ne_all, B, x, x, W))
    assert rc == 1 and b'64 MiB' in _lib.load().sawfem_last_error()
