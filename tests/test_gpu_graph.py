"""Graph-captured factorization (sawfem_mf_factor_graph): the captured launch sequence replayed
with new frequencies in the same buffer, interleaved with eager work, reproduces the eager
factorization bit for bit; a second chunk size gets its own graph; a kernel-selection knob
change drops the graphs; a refined solve through the graph path equals the eager one."""
import sys

import numpy as np
import pytest

from sawsim_amd import _lib, workloads
from sawsim_amd.piezo_fem import stream_ptr

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _factor(eng, fn, om, arena, B):
    _lib.call(fn, eng._plan, stream_ptr(), 0, B, om.data_ptr(), eng.beta, eng.Kaa.data_ptr(), eng.Maa.data_ptr(),
              eng._dptr(eng.Daa), arena.data_ptr(), 64)


def test_graph_replays_match_the_eager_factorization():
    _say('[test_graph_replays_match_the_eager_factorization]')
    cell = workloads.config3(mesh_um=0.12, points=6)
    eng = workloads.build_engine(cell)
    torch = eng.torch
    E = eng.mf.arena_elems
    f = cell.frequencies()
    buf = torch.empty(16, dtype=torch.float64, device=eng.dev)
    ga = torch.empty(E * 7, dtype=torch.complex64, device=eng.dev)
    ea = torch.empty(E * 7, dtype=torch.complex64, device=eng.dev)
    x = torch.randn((7, 1, eng.n_active), dtype=torch.complex128, device=eng.dev)
    uarena = torch.empty(eng.mf.uarena_elems * 7, dtype=torch.complex128, device=eng.dev)
    nodes = None
    for rep, shift in enumerate((0.0, 1.3e6, -2.1e6, 0.4e6)):
        om = torch.as_tensor(2 * np.pi * (np.linspace(f[0], f[-1], 7) + shift), device=eng.dev)
        buf[:7].copy_(om)
        _factor(eng, 'sawfem_mf_factor_graph', buf, ga, 7)
        _factor(eng, 'sawfem_mf_factor', om, ea, 7)
        # eager work between replays: a solve with the eager factors and some torch work
        _lib.call('sawfem_mf_solve', eng._plan, stream_ptr(), 0, 7, 1, ea.data_ptr(), uarena.data_ptr(), x.data_ptr(),
                  0, 64, None, 0)
        x.mul_(0.5)
        torch.cuda.synchronize()
        assert torch.equal(ga, ea), rep
        n_ = _lib.call_value('sawfem_mf_graph_nodes', eng._plan)
        assert n_ > 10 and (nodes is None or n_ == nodes)
        nodes = n_
    _say('graph nodes per factorization', nodes)
    # a second chunk size: its own graph; then the first size again (a replay)
    om5 = torch.as_tensor(2 * np.pi * np.linspace(f[0], f[-1], 5), device=eng.dev)
    buf[:5].copy_(om5)
    _factor(eng, 'sawfem_mf_factor_graph', buf, ga, 5)
    _factor(eng, 'sawfem_mf_factor', om5, ea, 5)
    torch.cuda.synchronize()
    assert torch.equal(ga[:E * 5], ea[:E * 5])
    om7 = torch.as_tensor(2 * np.pi * np.linspace(f[0], f[-1], 7), device=eng.dev)
    buf[:7].copy_(om7)
    _factor(eng, 'sawfem_mf_factor_graph', buf, ga, 7)
    _factor(eng, 'sawfem_mf_factor', om7, ea, 7)
    torch.cuda.synchronize()
    assert torch.equal(ga, ea)
    # a knob change drops the graphs: the next graph factorization follows the new setting
    eng.set_panel_lu(0)
    _factor(eng, 'sawfem_mf_factor_graph', buf, ga, 7)
    _factor(eng, 'sawfem_mf_factor', om7, ea, 7)
    torch.cuda.synchronize()
    assert torch.equal(ga, ea)


def test_refined_solve_through_the_graph_equals_the_eager_path():
    _say('[test_refined_solve_through_the_graph_equals_the_eager_path]')
    cell = workloads.config3(mesh_um=0.12, points=9)
    eng = workloads.build_engine(cell, gradient=True)
    f = cell.frequencies()
    eng.set_share(3)
    eng.use_graph = False
    Qe, dYe, _ = eng.sweep_charges(f, precision='mixed', gradient=True)
    eng.use_graph = True
    for _ in range(2):
        Qg, dYg, _ = eng.sweep_charges(f, precision='mixed', gradient=True)
        assert np.array_equal(Qe, Qg) and np.array_equal(dYe, dYg)
