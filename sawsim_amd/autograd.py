"""Y11(p) as a differentiable PyTorch function: the hand-written adjoint behind ``torch.autograd``.

    model = SweepModel(workloads.config3())              # or a Hex27 slab cell
    p = torch.tensor(model.params, requires_grad=True)
    y = model(p)                                         # complex128 Y11 at the cell's frequencies
    loss(y).backward()                                   # p.grad from the GPU adjoint

Forward: the piezo record is rebuilt from ``p``, K/M/D are re-assembled on the GPU and gathered
into the engine's active system (the symbolic analysis and the numeric plan are kept), then one
batched sweep returns Y11 and dY11/dp (paired or transposed adjoint, see ``FrequencySolver``).
Backward: for real parameters and complex outputs PyTorch propagates grad_p = Re(Jᴴ grad_y) with
J = dY11/dp, which is what ``_Y11.backward`` returns.  The reference has no gradient; this is the
north star's differentiable wrapper around the adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

from . import workloads


class SweepModel:
    """A cell, its GPU engine with the gradient context, and the parameter vector it was built from."""

    def __init__(self, cell, precision='mixed', V=1.0, batch=None, **engine_kw):
        if cell.params is None:
            raise ValueError('this cell has no design parameters')
        self.cell = cell
        self.precision = precision
        self.V = float(V)
        self.batch = batch
        self.hex = isinstance(cell, workloads.HexCell)
        self.engine = workloads.build_engine(cell, gradient=True, **engine_kw)
        self.params = np.asarray(cell.params, dtype=float).copy()
        self.sign = 1j if self.hex else -1j          # Y = +i w Q / V on slabs, -i w Q / V in 2D

    def set_params(self, p):
        """Rebuild the operator and the gradient context for parameter vector ``p``."""
        p = np.asarray(p, dtype=float)
        if p.shape != self.params.shape:
            raise ValueError(f'expected {self.params.size} parameters, got {p.size}')
        if np.array_equal(p, self.params):
            return
        cell, eng = self.cell, self.engine
        omega_top = 2 * np.pi * cell.band_hz[1]
        if self.hex:
            mats = workloads.hex_materials_for_params(cell, p)
            K, M, D = workloads.assemble_hex_cell(cell, mats)
            eng.update_operator(K, M, D, beta=p[15])
            mesh = cell.mesh
            eng.enable_gradient(mesh.hexas, mesh.pos_m, cell.material_list(mats), cell.pml_list(mats), omega_top, p)
        else:
            K, M, D = workloads.assemble_cell(cell, workloads.materials_for_params(cell, p))
            eng.update_operator(K, M, D, beta=p[15])
            eng.enable_gradient(cell.model, cell.mat_idx, cell.pml, omega_top, p, 'xy')
        self.params = p.copy()

    def __call__(self, params, freqs=None):
        f = self.cell.frequencies() if freqs is None else np.asarray(freqs, dtype=float)
        return _Y11.apply(params, self, f)


class _Y11(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, model, freqs):
        model.set_params(params.detach().cpu().numpy())
        Q, dY, _ = model.engine.sweep_charges(freqs, model.V, precision=model.precision, batch=model.batch,
                                              gradient=True)
        y = model.sign * 2 * np.pi * freqs * Q / model.V
        ctx.save_for_backward(torch.as_tensor(dY, dtype=torch.complex128))
        ctx.param_dtype = params.dtype
        return torch.as_tensor(y, dtype=torch.complex128, device=params.device)

    @staticmethod
    def backward(ctx, grad_y):
        (dY,) = ctx.saved_tensors
        g = (dY.conj().transpose(0, 1).to(grad_y.device) @ grad_y.to(torch.complex128)).real
        return g.to(ctx.param_dtype), None, None


def y11(params, model: SweepModel, freqs=None):
    """Functional form of :class:`SweepModel` ``__call__``."""
    return model(params, freqs)
