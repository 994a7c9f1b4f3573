"""GPU assembly of the coupled displacement-potential K and M (drop-in for
SawSim ``saw2d/piezo_fem.py``).

Same entry points and argument meaning as the reference:

* ``Model(pos_m, quads, elem_tag, N)`` (``piezo_fem.py:161-177``),
* ``PMLRegion`` (``piezo_fem.py:192-208``), ``material_index`` (``:211-219``),
* ``assemble(model, materials, elem_mat_idx, pml_regions, omega_top) -> (K, M)``
  (``:283-367``, ``[ux|uy|uz|phi]``, 4N unknowns),
* ``assemble_ME0(...) -> (K, M)`` (``:373-405``, ``[ux|uy|phi]``, 3N unknowns).

K and M come back as :class:`DeviceCSR` matrices resident in HBM on the structural
pattern of :class:`sawsim_amd.layout.KLayout`; ``.to_scipy()`` copies one to the host.
The element contractions run in ``sawfem_elem_q9`` and the scatter in the atomic-free
gather ``sawfem_gather_kmd``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .layout import KLayout
from .materials import Material, kernel_tensors, lift_projected

SCALE = 6.173191780858575e-11


@dataclass
class Model:
    pos_m: np.ndarray
    quads: np.ndarray
    elem_tag: np.ndarray
    N: int

    @property
    def ne(self) -> int:
        return int(self.quads.shape[0])

    def nodes_with_tag(self, line_tag: int, lines: np.ndarray) -> np.ndarray:
        sel = lines[lines[:, 3] == line_tag][:, :3].astype(np.int64) - 1
        return np.unique(sel)


@dataclass
class PMLRegion:
    tag: int
    kind: str
    dmax: float = 1e11
    n: int = 3
    xa: float = 0.0
    xp: float = 0.0
    ya: float = 0.0
    yp: float = 0.0


def material_index(model: Model, tag_map: dict) -> np.ndarray:
    idx = np.full(model.ne, -1, dtype=np.int64)
    for tag, mi in tag_map.items():
        idx[model.elem_tag == tag] = mi
    if (idx < 0).any():
        raise ValueError(f'element tags missing from the material map: {sorted(set(model.elem_tag[idx < 0].tolist()))}')
    return idx


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.SawfemError('sawsim_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch


def stream_ptr():
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


# Optional timing of the assembly and Bloch-gather launches: when set to a list, every such launch
# appends (kernel, start event, stop event, algorithmic HBM bytes) recorded on the launch stream.
KERNEL_LOG = None


def timed_call(kernel, nbytes, name, *args):
    """``_lib.call(name, *args)``, bracketed by events into KERNEL_LOG when it is a list."""
    if KERNEL_LOG is None:
        return _lib.call(name, *args)
    torch = _torch()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    rc = _lib.call(name, *args)
    b.record()
    KERNEL_LOG.append((kernel, a, b, int(nbytes)))
    return rc


class DeviceCSR:
    """Complex128 values in HBM on a shared structural pattern."""

    def __init__(self, layout: KLayout, values):
        self.layout = layout
        self.values = values
        self.shape = (layout.gdof, layout.gdof)
        self.node_xy = None

    @property
    def nnz(self):
        return self.layout.nnz

    def to_scipy(self):
        return self.layout.to_scipy(self.values.cpu().numpy())


class _Geometry:
    """Device copies of the mesh and the K/M gather lists, cached per model and layout."""

    def __init__(self, model: Model, nd: int):
        torch = _torch()
        dev = torch.device('cuda')
        self.layout = KLayout(model.quads, model.N, nd)
        kl = self.layout
        self.conn = torch.as_tensor(np.ascontiguousarray(model.quads, dtype=np.int32), device=dev)
        self.xy = torch.as_tensor(np.ascontiguousarray(model.pos_m[:, :2], dtype=np.float64), device=dev)
        self.tag = torch.as_tensor(np.ascontiguousarray(model.elem_tag, dtype=np.int32), device=dev)
        self.pair_row = torch.as_tensor(kl.pair_row.astype(np.int32), device=dev)
        self.node_ptr = torch.as_tensor(kl.node_ptr.astype(np.int64), device=dev)
        self.gptr = torch.as_tensor(kl.pair_gather.ptr, device=dev)
        # gather codes e*81 + i*9 + j from the flat (e, i*9 + j) element-entry index
        self.gidx = torch.as_tensor(kl.pair_gather.idx, device=dev)


_GEOM_CACHE = {}


def _geometry(model: Model, nd: int) -> _Geometry:
    key = (id(model), nd)
    g = _GEOM_CACHE.get(key)
    if g is None or g.layout.N != model.N or g.layout.ne != model.ne:
        g = _Geometry(model, nd)
        _GEOM_CACHE.clear()
        _GEOM_CACHE[key] = g
    return g


def _pml_arrays(pml_regions):
    if len(pml_regions) > 8:
        raise ValueError('at most 8 PML regions')
    tags = np.array([r.tag for r in pml_regions] or [0], dtype=np.int32)
    kinds = np.array([(1 if 'x' in r.kind else 0) | (2 if 'y' in r.kind else 0) for r in pml_regions] or [0],
                     dtype=np.int32)
    par = np.array([[r.dmax, r.n, r.xa, r.xp, r.ya, r.yp] for r in pml_regions] or [[0] * 6], dtype=np.float64)
    return tags, kinds, np.ascontiguousarray(par)


def assemble_device(model: Model, materials, elem_mat_idx, pml_regions, omega_top: float, nd: int,
                    damping_mask=None):
    """Run K1 + K2 on the GPU; returns (K, M[, D]) DeviceCSR on the structural pattern.

    ``damping_mask`` (bool per material) additionally gathers D = the K_uu part of the
    flagged materials (the reference's second assembly with e = eps = 0, ``solver.py:259-265``).
    """
    torch = _torch()
    dev = torch.device('cuda')
    g = _geometry(model, nd)
    kl = g.layout
    ne = model.ne
    C9, E9, EP, rho = kernel_tensors(materials)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    dC9, dE9, dEP, drho = t(C9), t(E9), t(EP), t(rho)
    midx = t(np.asarray(elem_mat_idx, dtype=np.int32))
    Ke = torch.empty(ne * nd * nd * 81, dtype=torch.complex128, device=dev)
    Me = torch.empty(ne * 81, dtype=torch.complex128, device=dev)
    tags, kinds, par = _pml_arrays(pml_regions)
    s = stream_ptr()
    # K1 writes the element matrices; it reads 9 node indices and coordinates per element
    timed_call('k_elem_q9', ne * (16 * (nd * nd + 1) * 81 + 9 * (4 + 16) + 8), 'sawfem_elem_q9', s, ne, nd - 1,
               g.conn.data_ptr(), g.xy.data_ptr(), midx.data_ptr(), g.tag.data_ptr(),
               dC9.data_ptr(), dE9.data_ptr(), dEP.data_ptr(), drho.data_ptr(), len(pml_regions),
               tags.ctypes.data, kinds.ctypes.data, par.ctypes.data, float(omega_top), SCALE,
               Ke.data_ptr(), Me.data_ptr())
    K = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
    M = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
    D = None
    dmask = torch.zeros(ne, dtype=torch.uint8, device=dev)
    if damping_mask is not None:
        D = torch.empty(kl.nnz, dtype=torch.complex128, device=dev)
        dm = np.asarray(damping_mask, dtype=bool)[np.asarray(elem_mat_idx)]
        dmask = t(dm.astype(np.uint8))
    # K2 reads every element entry once (Ke, Me) and its code, writes K, M (and D)
    nbytes = Ke.numel() * 16 + Me.numel() * 16 + g.gidx.numel() * 8 + kl.nnz * 16 * (3 if D is not None else 2)
    timed_call('k_gather_kmd', nbytes, 'sawfem_gather_kmd', s, kl.nnz_node, nd, nd - 1, g.pair_row.data_ptr(),
               g.node_ptr.data_ptr(), g.gptr.data_ptr(), g.gidx.data_ptr(), dmask.data_ptr(), Ke.data_ptr(),
               Me.data_ptr(), K.data_ptr(), M.data_ptr(), D.data_ptr() if D is not None else None)
    out = (DeviceCSR(kl, K), DeviceCSR(kl, M))
    if D is not None:
        out = out + (DeviceCSR(kl, D),)
    for mat in out:
        mat.node_xy = np.asarray(model.pos_m)[:, :2]
    return out


def assemble(model: Model, materials, elem_mat_idx, pml_regions, omega_top: float):
    """K, M (4N x 4N) for full 6x6 / 3x6 / 3x3 tensors, ``[ux|uy|uz|phi]``."""
    return assemble_device(model, materials, elem_mat_idx, pml_regions, omega_top, 4)


def assemble_ME0(model: Model, materials, elem_mat_idx, pml_regions, omega_top: float):
    """K, M (3N x 3N) for xy-projected tensors (C 3x3, e 2x3, eps 2x2), ``[ux|uy|phi]``."""
    lifted = []
    for m in materials:
        if np.shape(m.C) != (3, 3) or np.shape(m.e) != (2, 3) or np.shape(m.eps) != (2, 2):
            raise ValueError('ME0 requires projected xy tensors: C3x3, e2x3, eps2x2')
        lifted.append(lift_projected(m))
    return assemble_device(model, lifted, elem_mat_idx, pml_regions, omega_top, 3)
