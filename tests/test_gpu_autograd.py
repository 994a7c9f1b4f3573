"""The adjoint through torch.autograd: gradients of a real loss of Y11(p) against central
differences of the FP64 GPU sweep (plane config-3 cell and the coarse Hex27 slab)."""
import sys

import numpy as np
import pytest
import torch

from sawsim_amd import workloads
from sawsim_amd.autograd import SweepModel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _loss(y):
    return (y.abs() ** 2).sum() * 1e6 + (y.real * torch.arange(1, y.numel() + 1, dtype=torch.float64)).sum() * 1e3


CELLS = {
    'config3_coarse': lambda: workloads.config3(mesh_um=0.12, points=3),
    'config4_coarse': lambda: workloads.config4(mesh_um=0.2, points=3, beta_dk=1e-13, eta_eps=1e-3),
}


@pytest.mark.parametrize('name', list(CELLS))
def test_autograd_matches_central_differences(name):
    _say(f'[test_autograd_matches_central_differences] {name}')
    cell = CELLS[name]()
    model = SweepModel(cell, precision='mixed')
    p0 = model.params.copy()
    p = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    y = model(p)
    L = _loss(y)
    L.backward()
    grad = p.grad.numpy()
    ref = SweepModel(cell, precision='fp64')
    ks = [0, 6, 10, 13, 15, 16] + ([17] if p0.size == 18 else [])
    scale = np.where(p0 != 0, np.abs(p0), 1.0)       # angles (0 here) in degrees
    fds = {}
    for k in ks:
        h = 1e-6 * abs(p0[k]) if p0[k] != 0 else 1e-6 * 180 / np.pi
        vals = []
        for sgn in (1, -1):
            q = p0.copy()
            q[k] += sgn * h
            vals.append(float(_loss(ref(torch.tensor(q)))))
        fds[k] = (vals[0] - vals[1]) / (2 * h)
        _say(f'  p[{k}]: autograd {grad[k]:.8e} fd {fds[k]:.8e}')
    top = max(abs(fds[k]) * scale[k] for k in ks)
    for k in ks:
        assert abs(grad[k] - fds[k]) * scale[k] <= 1e-5 * top
    # the model's operator follows the parameters: back at p0 the forward is unchanged
    y2 = model(torch.tensor(p0))
    assert torch.allclose(y2, y.detach(), rtol=1e-9, atol=0)
