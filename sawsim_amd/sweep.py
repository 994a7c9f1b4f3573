"""GPU frequency sweep: Bloch reduction, batched multifrontal LU, refinement and Y11
(drop-in for SawSim ``saw2d/sweep.py``).

Same entry points as the reference:

* ``bc_periodic(model, lines, ..., ndisp)`` -> ``(fix, zheng, fu, merge_groups)``
  (``sweep.py:41-72``),
* ``FrequencySolver(K, M, fix, zheng, fu, merge_groups)`` with ``set_damping(D, beta)``,
  ``solve(omega, V) -> (disp, Q, |Y|)`` and ``sweep(fre, V) -> (Q, |Y|)``
  (``sweep.py:275-393``).

What runs where:

* host, once per cell: union-find merge, active set, gather lists, nested-dissection
  symbolic analysis (``layout.py``, ``symbolic.py``, ``plan.py``);
* GPU: Bloch gathers of K/M/D into the active system, RHS and charge weights (K3), the
  batched multifrontal factorization of ``A(w) = K + i w beta D - w^2 M`` for a chunk of
  frequencies, frontal triangular solves, FP64 residual + refinement and the charge
  reduction ``Q = kts . disp`` (``sweep.py:342-363``).

Precision modes:

* ``'fp64'``: complex128 fronts (plain path), one refinement step;
* ``'mixed'``: complex64 fronts with FP32 MFMA accumulation, refined against the FP64
  operator until the estimated relative error of Q is below ``ir_tol`` (see
  :func:`ir_error`); frequencies that stall are refactored in complex128.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .layout import BlochLayout, residual_groups
from .piezo_fem import DeviceCSR, Model, _torch, stream_ptr, timed_call
from .symbolic import analyse


# in the forward passes, fronts up to this order are solved by one workgroup (pivot triangle +
# panel) per frequency (measured on config 2 with the size buckets of a level on concurrent
# streams: 256 is the best of 128/192/256/320/448/640)
SOLVE_FULL_M = 256
# consecutive frequencies per complex64 factorization (FrequencySolver.solve_batch)
SHARE = 1
# 32-wide pivot panels of size buckets up to this front order are factored by one kernel per
# panel (diagonal block and both panel solves in one wave per front and frequency)
PANEL_LU_M = 256
# one-workgroup solve buckets whose largest factor block fits this many bytes are solved with
# the block staged in LDS by persistent workgroups (forward and backward); off by default:
# on config 2 it shortens the level-0 solves (193 -> 152 ms of kernel time per sweep) but the
# persistent grids crowd out the concurrent buckets (291 vs 294 pts/s); the tuner may pick it
TRI_LDS_BYTES = 0

# device memory every workspace request leaves free for the HIP runtime (queues, code objects,
# kernel scratch backing) and the caller's own tensors
HBM_RESERVE_BYTES = 16 * 10 ** 9


def hbm_available(keep=0):
    """Device bytes a new workspace may take: free memory, plus what torch caches unused, plus
    ``keep`` (a workspace that will be released for the new one), minus ``HBM_RESERVE_BYTES``."""
    torch = _torch()
    free, _ = torch.cuda.mem_get_info()
    cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
    return int(free + cached + keep - HBM_RESERVE_BYTES)


def check_workspace(nbytes, keep=0, what='front workspace'):
    """Raise ValueError (before anything is launched) when ``nbytes`` would break the reserve."""
    avail = hbm_available(keep)
    if nbytes > avail:
        raise ValueError(f'{what} of {nbytes / 1e9:.1f} GB exceeds the {max(avail, 0) / 1e9:.1f} GB of device memory '
                         f'available above the {HBM_RESERVE_BYTES / 1e9:.0f} GB reserve; use fewer frequencies per chunk')


def ir_error(dq, dq_prev):
    """Estimated relative error of a refined quantity after its latest correction.

    ``dq`` is the relative size of the latest correction, ``dq_prev`` that of the one before
    (inf after the first).  Refinement converges linearly; with the observed rate
    r = dq / dq_prev below 1/2, the corrections still to come sum to about dq r / (1 - r)
    (geometric series).  Otherwise (first step, slow or stalled contraction) the estimate is
    dq itself.  This is the reported estimate; the refinement stops on dq itself (see
    ``FrequencySolver._refine``), because the rate can change: complex64 factors can remove the
    dominant error in one step and then converge slowly on a component they resolve poorly
    (seen on a design cell near resonance: corrections 4e-4, 2.4e-6, 8.6e-8, 9.3e-8 with a true
    error of 7e-7 after the third), which an extrapolation from the fast step hides.
    """
    dq = np.asarray(dq, dtype=float)
    prev = np.asarray(dq_prev, dtype=float)
    with np.errstate(divide='ignore', invalid='ignore'):
        r = np.where(np.isfinite(prev), dq / np.maximum(prev, 1e-300), 1.0)
        return np.where(r < 0.5, dq * r / (1.0 - r), dq)


def bc_periodic(model: Model, lines, tag_bottom=21, tag_left=19, tag_right=20, tag_zheng=17, tag_fu=18, ndisp=3):
    """Bloch-periodic (zero phase) boundary sets of an SP cell (``sweep.py:41-72``).

    Left/right nodes without the bottom corners are paired by ascending y; the pairs
    merge all displacement components and the potential, the bottom corners merge the
    potential only; bottom displacements are fixed.
    """
    N = model.N
    nodes = lambda tag: model.nodes_with_tag(tag, lines)
    bottom, left, right = nodes(tag_bottom), nodes(tag_left), nodes(tag_right)
    zheng, fu = nodes(tag_zheng), nodes(tag_fu)
    cl, cr = np.intersect1d(bottom, left), np.intersect1d(bottom, right)
    L = np.setdiff1d(left, cl)
    R = np.setdiff1d(right, cr)
    L = L[np.argsort(model.pos_m[L, 1], kind='stable')]
    R = R[np.argsort(model.pos_m[R, 1], kind='stable')]
    if L.size != R.size:
        raise RuntimeError('left/right periodic node counts differ')
    masters = np.concatenate([L + a * N for a in range(ndisp + 1)] + [cl + ndisp * N])
    slaves = np.concatenate([R + a * N for a in range(ndisp + 1)] + [cr + ndisp * N])
    groups = [np.array([a, b]) for a, b in zip(masters, slaves)]
    fix = np.concatenate([bottom + a * N for a in range(ndisp)])
    return fix, zheng + ndisp * N, fu + ndisp * N, groups


class _ResidualOp:
    """FP64 residual r = rhs - A(w) x of a chunk (A^T with a transposed-slot map), through the
    frequency-major layout: x and rhs are transposed to [n][Bp] (Bp = B padded to 64), the
    residual kernel reads each A_aa entry once for all frequencies and each x row once per
    row group (the active unknowns of one merged node share their column list), r is
    transposed back.  With ``ld`` the operands are already row-major [n][ld] (the grouped
    refinement keeps its vectors so) and nothing is transposed.

    A module-level class on purpose: the [n][Bp] buffers must be freed when the chunk's solve
    returns (a class defined inside a method sits in a reference cycle, and its buffers would
    wait for Python's cyclic garbage collector)."""

    def __init__(self, solver, B, om, ld=0):
        torch = solver.torch
        self.solver, self.B, self.om, self.ld = solver, B, om, ld
        n = solver.n_active
        self.Bp = -(-B // 64) * 64
        if ld:                 # x, rhs and r already row-major [n][ld]: no staging buffers
            self.xt = self.rt = None
            return
        self.xt = torch.empty((n, self.Bp), dtype=torch.complex128, device=solver.dev)
        self.rt = torch.empty((n, self.Bp), dtype=torch.complex128, device=solver.dev)

    # below this fraction of active vectors a step gathers them (compact residual)
    COMPACT = 0.9

    def to_t(self, v):
        """The right-hand sides ``v`` ([B][n]) transposed for the residual kernel; ``v`` itself is
        kept for the compact steps, which add it while scattering back."""
        sv = self.solver
        n = sv.n_active
        self.rhs = v
        out = sv.torch.empty((n, self.Bp), dtype=sv.torch.complex128, device=sv.dev)
        _lib.call('sawfem_transpose', stream_ptr(), self.B, n, n, self.Bp, v.data_ptr(), out.data_ptr())
        return out

    def __call__(self, x, rhst, r, tslot, active=None, per=1):
        """r = rhs - A x.  ``active`` (bool per frequency, ``per`` vectors each): when fewer than
        COMPACT of them are active, only their vectors are gathered, evaluated and scattered back
        (the rows of the other vectors of r are left as they are).  Row-major operands (``ld``)
        are read and written in place, the active vectors through a column map."""
        sv = self.solver
        n, B, Bp = sv.n_active, self.B, self.Bp
        s = stream_ptr()
        gptr, grows, ng = sv.res_groups
        if self.ld:
            cmap, nc = None, B
            if active is not None and active.sum() * per < self.COMPACT * B:
                vec = (np.flatnonzero(active)[:, None] * per + np.arange(per)).ravel().astype(np.int32)
                if vec.size == 0:
                    return
                cmap, nc = sv.torch.as_tensor(vec, device=sv.dev), int(vec.size)
            _lib.call('sawfem_residual_sv', s, ng, nc, self.ld, None if cmap is None else cmap.data_ptr(),
                      gptr.data_ptr(), grows.data_ptr(), sv.aa_indptr.data_ptr(), sv.aa_indices.data_ptr(),
                      sv._dptr(tslot), sv.Kaa.data_ptr(), sv.Maa.data_ptr(), sv._dptr(sv.Daa), self.om.data_ptr(),
                      sv.beta, x.data_ptr(), rhst.data_ptr(), r.data_ptr())
            return
        if active is not None and active.sum() * per < self.COMPACT * B:
            torch = sv.torch
            vec = (np.flatnonzero(active)[:, None] * per + np.arange(per)).ravel().astype(np.int32)
            nsel = int(vec.size)
            if nsel == 0:
                return
            sel = torch.as_tensor(vec, device=sv.dev)
            Bc = -(-nsel // 64) * 64
            xt = self.xt.view(-1)[:n * Bc].view(n, Bc)
            rt = self.rt.view(-1)[:n * Bc].view(n, Bc)
            om = self.om[sel.long()].contiguous()
            _lib.call('sawfem_transpose_map', s, nsel, n, n, Bc, x.data_ptr(), xt.data_ptr(), sel.data_ptr(), None,
                      None)
            _lib.call('sawfem_residual_sv', s, ng, nsel, Bc, None, gptr.data_ptr(), grows.data_ptr(),
                      sv.aa_indptr.data_ptr(), sv.aa_indices.data_ptr(), sv._dptr(tslot), sv.Kaa.data_ptr(),
                      sv.Maa.data_ptr(), sv._dptr(sv.Daa), om.data_ptr(), sv.beta, xt.data_ptr(), None,
                      rt.data_ptr())
            _lib.call('sawfem_transpose_map', s, n, nsel, Bc, n, rt.data_ptr(), r.data_ptr(), None, sel.data_ptr(),
                      self.rhs.data_ptr())
            return
        _lib.call('sawfem_transpose', s, B, n, n, Bp, x.data_ptr(), self.xt.data_ptr())
        _lib.call('sawfem_residual_sv', s, ng, B, Bp, None, gptr.data_ptr(), grows.data_ptr(), sv.aa_indptr.data_ptr(),
                  sv.aa_indices.data_ptr(), sv._dptr(tslot), sv.Kaa.data_ptr(), sv.Maa.data_ptr(),
                  sv._dptr(sv.Daa), self.om.data_ptr(), sv.beta, self.xt.data_ptr(), rhst.data_ptr(),
                  self.rt.data_ptr())
        _lib.call('sawfem_transpose', s, n, B, Bp, n, self.rt.data_ptr(), r.data_ptr())


class FrequencySolver:
    """Per-frequency solve of the Bloch-reduced piezoelectric system on one GPU."""

    n_params = 17          # dY11/dp parameters (params.NAMES)

    def __init__(self, K: DeviceCSR, M: DeviceCSR, fix, zheng, fu, merge_groups, *, node_xy=None, period=None,
                 leaf_weight=64):
        bl = BlochLayout(K.layout, fix, zheng, fu, merge_groups)
        xy = node_xy if node_xy is not None else getattr(K, 'node_xy', None)
        if xy is None:
            raise ValueError('node coordinates are needed for the nested-dissection ordering')
        xy = np.asarray(xy)
        if period is None and len(merge_groups):
            g = np.asarray(merge_groups[0]).ravel()
            period = float(abs(xy[g[1] % K.layout.N, 0] - xy[g[0] % K.layout.N, 0])) or None
        self._setup(K, M, bl, bl.node_of, xy, period, leaf_weight)

    def _setup(self, K, M, bl, sv_of, sv_xy, period, leaf_weight):
        """Device copies of the reduction maps, the active operators and the numeric plan.

        ``sv_of`` groups the active unknowns into supervariables (ids into ``sv_xy``, the
        coordinates the nested dissection cuts on); ``period`` is the Bloch period in x.
        """
        torch = _torch()
        self.torch = torch
        dev = self.dev = torch.device('cuda')
        kl = self.layout = K.layout
        if M.layout is not kl:
            raise ValueError('K and M must share one layout')
        self.gdof = kl.gdof
        self.bloch = bl
        self.zheng, self.fu = bl.zheng, bl.fu
        self.active, self.slaves, self.masters = bl.active, bl.slaves, bl.masters
        n = self.n_active = bl.n_active
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
        self._aa = (t(bl.aa.ptr), t(bl.aa.idx))
        self._rh = (t(bl.rhs.ptr), t(bl.rhs.idx))
        self._kt = (t(bl.kts.ptr), t(bl.kts.idx))
        self.aa_indptr = t(bl.aa_indptr.astype(np.int64))
        self.aa_indices = t(bl.aa_indices.astype(np.int32))
        gptr, grows = residual_groups(bl.aa_indptr, bl.aa_indices, bl.node_of)
        self.res_groups = (t(gptr), t(grows), int(gptr.size - 1))
        z = lambda k: torch.empty(k, dtype=torch.complex128, device=dev)
        self.Kaa, self.Maa, self.rK, self.rM = z(bl.nnz_aa), z(bl.nnz_aa), z(n), z(n)
        self.Daa = self.rD = None
        self.beta = 0.0
        self._kts = z(n + 1)
        self.update_operator(K, M)
        # symbolic analysis and numeric plan
        t0 = time.time()
        self.symbolic = analyse(bl.aa_indptr, bl.aa_indices, sv_of, bl.comp_of, np.asarray(sv_xy),
                                period=period, leaf_weight=leaf_weight)
        self.mf = self.symbolic.mf
        self.analysis_seconds = time.time() - t0
        self._plan = self._create_plan()
        self.share, self._group = SHARE, 0
        self.use_graph, self._om_buf = True, None
        self.set_solve_full(SOLVE_FULL_M)
        self.set_panel_lu(PANEL_LU_M)
        self.set_tri_lds(TRI_LDS_BYTES)
        self.stats = dict(self.symbolic.stats)

    def set_panel_lu(self, max_m):
        """Factor the 32-wide panels of size buckets up to front order ``max_m`` with one kernel
        per panel (``sawfem_mf_set_panel_lu``); 0 uses the diagonal-block kernel and two GEMMs."""
        return _lib.call_value('sawfem_mf_set_panel_lu', self._plan, int(max_m))

    def set_tri_lds(self, max_bytes):
        """Solve one-workgroup buckets whose factor block fits ``max_bytes`` from LDS
        (``sawfem_mf_set_tri_lds``); 0 keeps the pivot-triangle kernels."""
        return _lib.call_value('sawfem_mf_set_tri_lds', self._plan, int(max_bytes))

    def set_group_rc(self, rc_max):
        """At most ``rc_max`` vectors per workgroup in the solves of shared factors
        (``sawfem_mf_set_group_rc``; 1..16)."""
        return _lib.call_value('sawfem_mf_set_group_rc', self._plan, int(rc_max))

    def set_solve_full(self, max_m):
        """Solve fronts of order <= ``max_m`` with one workgroup each (``sawfem_mf_set_solve_full``)."""
        return _lib.call_value('sawfem_mf_set_solve_full', self._plan, int(max_m))

    # ---------------------------------------------------------------- plan
    def _create_plan(self):
        import ctypes as C
        mf = self.mf
        h = C.c_void_p()
        a = lambda x: np.ascontiguousarray(x).ctypes.data
        keep = [np.ascontiguousarray(x) for x in (mf.m, mf.p, mf.rows_ptr, mf.rows, mf.child_ptr, mf.child_ids,
                                                   mf.inv_ptr, mf.inv, mf.lvl_ptr, mf.lvl_fronts, mf.ent_front,
                                                   mf.ent_slot, mf.ent_loc, mf.ent_lvl_ptr)]
        _lib.call('sawfem_mf_create', C.byref(h), mf.n, mf.nf, *[k.ctypes.data for k in keep[:8]],
                  mf.lvl_ptr.size - 1, keep[8].ctypes.data, keep[9].ctypes.data, mf.ent_front.size,
                  keep[10].ctypes.data, keep[11].ctypes.data, keep[12].ctypes.data, keep[13].ctypes.data)
        return h

    def __del__(self):
        try:
            if getattr(self, '_plan', None) is not None:
                _lib.raw('sawfem_mf_destroy', self._plan)
                self._plan = None
        except Exception:
            pass

    def update_operator(self, K: DeviceCSR, M: DeviceCSR, D: DeviceCSR = None, beta: float = None):
        """Re-gather the active system from new K/M (and D) values on the same mesh and boundary
        sets; the symbolic analysis and the numeric plan are kept (design loops, finite differences)."""
        if K.layout.nnz != self.layout.nnz or M.layout.nnz != self.layout.nnz:
            raise ValueError('K/M layout differs from the one this solver was built for')
        bl = self.bloch
        n = self.n_active
        s = stream_ptr()
        # K3 reads 16 B per listed source entry and source, the 8 B index, and writes 16 B per
        # output and source
        g3 = lambda ptr, nout, nsrc: int(ptr[1].numel() * (16 * nsrc + 8) + 8 * (nout + 1) + 16 * nout * nsrc)
        timed_call('k_gather3', g3(self._aa, bl.nnz_aa, 2), 'sawfem_gather3', s, bl.nnz_aa, self._aa[0].data_ptr(),
                   self._aa[1].data_ptr(), K.values.data_ptr(), M.values.data_ptr(), None, self.Kaa.data_ptr(),
                   self.Maa.data_ptr(), None)
        timed_call('k_gather3', g3(self._rh, n, 2), 'sawfem_gather3', s, n, self._rh[0].data_ptr(),
                   self._rh[1].data_ptr(), K.values.data_ptr(), M.values.data_ptr(), None, self.rK.data_ptr(),
                   self.rM.data_ptr(), None)
        timed_call('k_gather3', g3(self._kt, n + 1, 1), 'sawfem_gather3', s, n + 1, self._kt[0].data_ptr(),
                   self._kt[1].data_ptr(), K.values.data_ptr(), None, None, self._kts.data_ptr(), None, None)
        self.kts_a = self._kts[:n]
        self.kts_z = complex(self._kts[n].item())
        self._sym = None
        if D is not None:
            self.set_damping(D, self.beta if beta is None else beta)

    def set_damping(self, D: DeviceCSR, beta: float):
        """Rayleigh stiffness damping A = K + i w beta D - w^2 M (``sweep.py:316-323``)."""
        torch = self.torch
        bl = self.bloch
        self.Daa = torch.empty(bl.nnz_aa, dtype=torch.complex128, device=self.dev)
        self.rD = torch.empty(self.n_active, dtype=torch.complex128, device=self.dev)
        s = stream_ptr()
        _lib.call('sawfem_gather3', s, bl.nnz_aa, self._aa[0].data_ptr(), self._aa[1].data_ptr(),
                  D.values.data_ptr(), None, None, self.Daa.data_ptr(), None, None)
        _lib.call('sawfem_gather3', s, self.n_active, self._rh[0].data_ptr(), self._rh[1].data_ptr(),
                  D.values.data_ptr(), None, None, self.rD.data_ptr(), None, None)
        self.beta = float(beta)
        self._sym = None

    # ---------------------------------------------------------------- gradient context
    def enable_gradient(self, model, mat_idx, pml_regions, omega_top, params, plane='xy', material=0):
        """Prepare dY11/dp for the 17 parameters of material ``material`` (see ``params.py``).

        ``model``/``mat_idx``/``pml_regions``/``omega_top`` are the assembly inputs; ``params`` the
        parameter vector the material tensors were built from.
        """
        from . import params as prm
        from .piezo_fem import _pml_arrays
        torch = self.torch
        dev = self.dev
        bl = self.bloch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
        elems = np.flatnonzero(np.asarray(mat_idx) == material).astype(np.int32)
        self._grad = dict(
            elems=t(elems), ne0=int(elems.size),
            conn=t(np.asarray(model.quads, dtype=np.int32)), xy=t(np.asarray(model.pos_m)[:, :2].astype(np.float64)),
            tag=t(np.asarray(model.elem_tag, dtype=np.int32)), pml=_pml_arrays(pml_regions), npml=len(pml_regions),
            omega_top=float(omega_top), N=int(model.N),
            pos_rep=t(bl.pos[bl.rep].astype(np.int32)),
            zmask=t(np.isin(np.arange(self.gdof), bl.zheng).astype(np.uint8)),
            params=np.asarray(params, dtype=float), plane=plane)
        self._grad['tensors'] = prm.lifted_tensors(params, plane)
        self._grad['jac'] = prm.jacobian(params, plane)
        from .piezo_fem import SCALE
        self._gradient_maps(SCALE)

    def _gradient_maps(self, scale):
        """Transposed-entry slots of the active pattern (transposed residual), the scaling
        D = diag(1 on displacements, 1/scale on the potential) of the paired adjoint."""
        torch = self.torch
        bl = self.bloch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=self.dev)
        n = self.n_active
        r = np.repeat(np.arange(n, dtype=np.int64), np.diff(bl.aa_indptr))
        ip = np.ascontiguousarray(bl.aa_indptr, dtype=np.int64)
        ix = np.ascontiguousarray(bl.aa_indices, dtype=np.int64)
        tslot = np.empty(ix.size, dtype=np.int64)
        _lib.call('sawfem_transpose_slots', n, ip.ctypes.data, ix.ctypes.data, tslot.ctypes.data)
        self.tslot = t(tslot)
        is_phi = bl.comp_of == (self.layout.nd - 1)
        self._dscale = t(np.where(is_phi, 1.0 / scale, 1.0))
        self._arow = t(r)
        self._sym = None

    def scaled_symmetric(self, rtol=1e-12):
        """True when S = D A_aa is symmetric for D = diag(1 on displacements, 1/SCALE on the
        potential), checked on K, M and D separately.  The piezoelectric blocks satisfy
        K_phi,u = SCALE K_u,phi^T, the Bloch merge and the Dirichlet eliminations keep the
        symmetry of S, so A^-T g = D A^-1 (D^-1 g): the adjoint solve becomes a second
        right-hand side of the forward solve."""
        if self._sym is None:
            d = self._dscale
            dr, dc = d[self._arow], d[self.aa_indices.long()]
            ok = True
            for A in (self.Kaa, self.Maa, self.Daa):
                if A is None:
                    continue
                S = A * dr
                St = A[self.tslot] * dc
                ok &= bool((S - St).abs().max() <= rtol * S.abs().max())
            self._sym = ok
        return self._sym

    def _param_gradient(self, W, omegas, V):
        """dY11/dp (B, 17) from the element sensitivity sums W (B, NT)."""
        from .piezo_fem import SCALE
        g = self._grad
        ndu = self.layout.nd - 1
        NR = 2 * ndu
        rows = [[0, 5], [8, 1], [7, 6]]
        R = [rows[i // 2][i % 2] for i in range(NR)]
        Wc = W[:, :NR * NR].reshape(-1, NR, NR)
        We = W[:, NR * NR:NR * NR + 2 * NR].reshape(-1, 2, NR)
        Wp = W[:, NR * NR + 2 * NR:].reshape(-1, 2, 2)
        dC, dE, dP = g['jac']
        C9 = g['tensors'][0]
        dCs = dC[:, R][:, :, R]                      # (17, NR, NR)
        dEs = dE[:, :2][:, :, R]                     # (17, 2, NR)
        dPs = dP[:, :2, :2]
        w = np.asarray(omegas)[:, None]
        fac = (1 + 1j * w * self.beta)
        dQ = (fac * SCALE * np.einsum('bij,kij->bk', Wc, dCs) + np.einsum('bij,kij->bk', We, dEs)
              - np.einsum('bij,kij->bk', Wp, dPs))
        dQ[:, 15] = (1j * w[:, 0]) * SCALE * np.einsum('bij,ij->b', Wc, C9[np.ix_(R, R)])
        return -1j * w / V * dQ

    # ---------------------------------------------------------------- numeric
    def arena_bytes(self, batch, precision):
        """Bytes of ``batch`` frequencies' fronts."""
        return self.mf.arena_elems * batch * (8 if precision == 'mixed' else 16)

    def chunk_arena_bytes(self, batch, precision, share=None):
        """Front bytes of a chunk of ``batch`` frequencies: one factorization per ``share``
        of them in mixed precision (default ``self.share``)."""
        share = max(1, int(self.share if share is None else share)) if precision == 'mixed' else 1
        return self.arena_bytes(-(-batch // share), precision)

    def _dptr(self, t):
        return None if t is None else t.data_ptr()

    def _arena(self, count, dtype):
        """Front arena of ``count`` entries, carved from one workspace kept across calls and
        grown on demand: chunks of either precision reuse the same HBM instead of leaving the
        allocator with cached 100 GB blocks of the wrong size."""
        torch = self.torch
        nbytes = count * (8 if dtype == torch.complex64 else 16)
        ws = getattr(self, '_ws', None)
        if ws is None or ws.numel() < nbytes:
            check_workspace(nbytes, keep=0 if ws is None else ws.numel())
            self._ws = None
            del ws
            torch.cuda.empty_cache()
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        return self._ws[:nbytes].view(dtype)

    def chunk_bytes(self, batch, precision='mixed', gradient=False, share=None):
        """Device bytes one chunk of ``batch`` frequencies takes at its peak, counted from what
        solve_batch allocates: front arena; update vectors (two right-hand sides with the
        gradient); right-hand sides, solutions, residuals and their temporaries ([B][R][n]); the
        three frequency-major residual buffers ([n][Bp], Bp = R B padded to 64); the element sums
        of the gradient contraction; 10 % on top of everything but the arena, whose size is exact.
        With ``share`` (default ``self.share``; mixed precision) the arena holds one
        factorization per ``share`` frequencies."""
        R = 2 if gradient else 1
        share = max(1, int(self.share if share is None else share)) if precision == 'mixed' else 1
        n = self.n_active
        Bp = -(-R * batch // 64) * 64
        vecs = (12 if gradient else 4) * batch + 3 * Bp
        work = 0
        if gradient and getattr(self, '_grad', None) is not None:
            nd = self.layout.nd - 1
            work = 16 * batch * self._grad['ne0'] * ((2 * nd) ** 2 + 4 * nd + 4)
        other = 16 * self.mf.uarena_elems * R * batch + 16 * n * vecs + work
        return int(self.arena_bytes(-(-batch // share), precision) + 1.1 * other)

    def _workspace(self, nbytes):
        """``nbytes`` of the kept workspace (grown like the front arena)."""
        return self._arena(-(-int(nbytes) // 8), self.torch.complex64).view(self.torch.uint8)[:int(nbytes)]

    def sweep_chunk_native(self, omegas, V=1.0, precision='mixed', ir_tol=1e-8, max_ir=8, fallback=True, share=1):
        """One chunk of frequencies through the C ABI alone (``sawfem_sweep_chunk_share``: factor,
        solve, refinement and the fallbacks in the library; ``share`` consecutive frequencies per
        complex64 factorization).  Returns (Q complex (B,), estimated relative error (B,), flags
        (B,): bit 0 refactored in complex128, bit 1 not converged, bit 2 own complex64 factors after
        sharing, status: 0, 5 non-finite charge or 6 some frequencies refactored)."""
        import ctypes as C
        om = np.ascontiguousarray(np.asarray(omegas, dtype=np.float64).ravel())
        B = om.size
        if B == 0:
            raise ValueError('sweep_chunk_native: no frequencies')
        if precision not in ('mixed', 'fp64'):
            raise ValueError("precision must be 'mixed' or 'fp64', not %r" % (precision,))
        prec = 0 if precision == 'mixed' else 1
        op = _lib.SweepOperator(self.n_active, self.aa_indptr.data_ptr(), self.aa_indices.data_ptr(),
                                self.Kaa.data_ptr(), self.Maa.data_ptr(), self._dptr(self.Daa), self.rK.data_ptr(),
                                self.rM.data_ptr(), self._dptr(self.rD), self.kts_a.data_ptr(), self.kts_z.real,
                                self.kts_z.imag, self.beta)
        nbytes = _lib.load().sawfem_sweep_chunk_share_bytes(self._plan, self.n_active, B, prec, int(share))
        if nbytes < 0:
            raise _lib.SawfemError('sawfem_sweep_chunk_share_bytes failed')
        ws = self._workspace(nbytes)
        Q = np.empty(2 * B)
        err = np.empty(B)
        flags = np.zeros(B, dtype=np.int32)
        rc = _lib.raw('sawfem_sweep_chunk_share', self._plan, stream_ptr(), C.byref(op), B, om.ctypes.data, float(V),
                      prec, int(share), float(ir_tol), int(max_ir), int(bool(fallback)), ws.data_ptr(), int(nbytes),
                      Q.ctypes.data, err.ctypes.data, flags.ctypes.data)
        if rc not in (_lib.OK, _lib.NON_FINITE, _lib.REFACTORED):
            msg = _lib.load().sawfem_last_error().decode(errors='replace')
            raise _lib.SawfemError(f'sawfem_sweep_chunk failed with status {rc}: {msg}')
        return Q.view(complex), err, flags, rc

    def release_workspace(self):
        """Free the arena workspace (it is otherwise kept for the next solve_batch)."""
        self._ws = None
        self.torch.cuda.empty_cache()

    def solve_batch(self, omegas, V=1.0, precision='mixed', ir_tol=1e-8, max_ir=8, nb=None, fallback=True,
                    gradient=False, adjoint='auto', share=None):
        """Solve all ``omegas`` in one batched factorization.

        Returns dict with ``x`` (B, n) complex128 device tensor, ``Q`` (B,) complex numpy,
        ``ir_steps``, ``dq`` (last relative change of Q per frequency), ``err`` (its
        estimated remaining error), ``refactored`` (complex128 refactorization used),
        ``own_factors`` (refactored in complex64 after sharing failed); with ``gradient`` also
        ``dY`` and the adjoint's ``adj_*`` entries.  ``adjoint``: 'auto' pairs the adjoint with
        the forward solve when :meth:`scaled_symmetric` holds, 'paired' requires it,
        'transposed' always uses separate transposed solves.

        ``share`` (mixed precision; default ``self.share``): consecutive frequencies per
        factorization.  Groups of ``share`` frequencies are factored once, at the mean of their
        angular frequencies, and each frequency is refined against its own FP64 operator with
        those factors (``sawfem_mf_solve_shared``): a nearby frequency's factors precondition
        the refinement, which takes a few more steps instead of a factorization.  Frequencies
        that do not converge that way are refactored with their own complex64 factors, and
        those that still do not, in complex128.
        """
        torch = self.torch
        dev = self.dev
        omegas = np.asarray(omegas, dtype=np.float64).ravel()
        B, n = omegas.size, self.n_active
        if B == 0:
            raise ValueError('solve_batch: no frequencies')
        if precision not in ('mixed', 'fp64'):
            raise ValueError("precision must be 'mixed' or 'fp64', not %r" % (precision,))
        prec = 0 if precision == 'mixed' else 1
        nb = nb or (64 if prec == 0 else 32)
        paired = False
        if gradient and adjoint != 'transposed':
            paired = self.scaled_symmetric()
            if adjoint == 'paired' and not paired:
                raise ValueError('D A is not symmetric: the paired adjoint does not apply')
        R = 2 if paired else 1
        share = max(1, int(self.share if share is None else share)) if prec == 0 else 1
        share = min(share, B)
        s = stream_ptr()
        om = torch.as_tensor(omegas, device=dev)
        group = np.arange(B) // share
        Bf = int(group[-1]) + 1
        if share > 1:
            om_f = torch.as_tensor(np.bincount(group, weights=omegas) / np.bincount(group), device=dev)
        else:
            om_f = om
        self._group = share if share > 1 else 0
        arena = self._arena(self.mf.arena_elems * Bf, torch.complex64 if prec == 0 else torch.complex128)
        uarena = torch.empty(max(1, self.mf.uarena_elems * B * R), dtype=torch.complex128, device=dev)
        self._factor(prec, Bf, om_f, arena, nb)
        rhs = torch.empty((B, n), dtype=torch.complex128, device=dev)
        _lib.call('sawfem_rhs', s, n, B, om.data_ptr(), self.beta, float(V), self.rK.data_ptr(), self.rM.data_ptr(),
                  self._dptr(self.rD), rhs.data_ptr())
        if paired:
            out = self._solve_paired(om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir)
        else:
            out = self._solve_single(om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir)
            if gradient:
                out.update(self._adjoint(om, omegas, V, prec, nb, arena, uarena, rhs, out['x'], ir_tol, max_ir))
        self._group = 0
        x, err = out['x'], out['err']
        out['perturbed'], out['pivot_min_ratio'] = self.pivot_stats(reset=True)
        out['share'], out['factorizations'] = share, Bf
        refactored = np.zeros(B, dtype=bool)
        own = np.zeros(B, dtype=bool)
        if prec == 0 and fallback:
            conv = out['converged'] & out.get('adj_converged', True)
            bad = np.flatnonzero(~conv)
            if bad.size:
                del arena, uarena
                # shared factors: their own complex64 factorization first (which falls back to
                # complex128 itself); own factors: complex128.  In chunks that fit the workspace.
                sub_prec = 'mixed' if share > 1 else 'fp64'
                nfit = max(1, int(self._ws.numel() // self.arena_bytes(1, sub_prec)))
                for c0 in range(0, bad.size, nfit):
                    part = bad[c0:c0 + nfit]
                    sub = self.solve_batch(omegas[part], V, precision=sub_prec, ir_tol=ir_tol, max_ir=max_ir,
                                           gradient=gradient, adjoint=adjoint, share=1)
                    if sub_prec == 'mixed':
                        own[part] = True
                        refactored[part] = sub['refactored']
                        out['factorizations'] += sub['factorizations']
                    tb = self.torch.as_tensor(part, device=self.dev)
                    x[tb] = sub['x']
                    if 'lam' in out:
                        out['lam'][tb] = sub['lam']
                    for key in ('Q', 'dq', 'err', 'dY', 'adj_dq', 'adj_err', 'converged', 'adj_converged', 'dwr_mask'):
                        if key in out:
                            out[key][part] = sub[key]
                    out['perturbed_' + sub_prec] = out.get('perturbed_' + sub_prec, 0) + sub['perturbed']
                    del sub
                if sub_prec == 'fp64':
                    refactored[bad] = True
        out['refactored'] = refactored
        out['own_factors'] = own
        return out

    def _factor(self, prec, Bf, om_f, arena, nb):
        """Factor the chunk: with ``use_graph`` one HIP graph launch (``sawfem_mf_factor_graph``),
        the frequencies copied into a buffer the engine keeps, so every chunk of the same size
        (and workspace) replays one captured graph; else launch by launch (``sawfem_mf_factor``)."""
        s = stream_ptr()
        if not self.use_graph:
            _lib.call('sawfem_mf_factor', self._plan, s, prec, Bf, om_f.data_ptr(), self.beta, self.Kaa.data_ptr(),
                      self.Maa.data_ptr(), self._dptr(self.Daa), arena.data_ptr(), nb)
            return
        if self._om_buf is None or self._om_buf.numel() < Bf:
            self._om_buf = self.torch.empty(max(Bf, 64), dtype=self.torch.float64, device=self.dev)
        self._om_buf[:Bf].copy_(om_f)
        _lib.call('sawfem_mf_factor_graph', self._plan, s, prec, Bf, self._om_buf.data_ptr(), self.beta,
                  self.Kaa.data_ptr(), self.Maa.data_ptr(), self._dptr(self.Daa), arena.data_ptr(), nb)

    def set_share(self, k):
        """Consecutive frequencies per complex64 factorization (see :meth:`solve_batch`)."""
        k = int(k)
        if k < 1:
            raise ValueError('share must be >= 1')
        self.share = k
        return k

    def _solve(self, prec, B, R, arena, uarena, vec, transpose, nb, active=None, ld=0):
        """Forward + backward solve of the B x R right-hand sides ``vec`` in place with the chunk's
        factors; ``active`` (bool [B], or None for all) selects the frequencies to solve.

        When ``solve_batch`` factored groups of G frequencies (complex64), the vectors of a group
        are solved together (``sawfem_mf_solve_groups``): a group is solved when any of its
        frequencies is active, and the others of the group must hold zero right-hand sides
        (their solution is then zero).  ``ld``: ``vec`` is row-major [n][ld] (grouped solves only,
        ``sawfem_mf_solve_groups_rows``), else [B][R][n]."""
        G = self._group
        idx = None
        if G > 1:
            nfull = B // G
            tail = 1
            if active is not None:
                ga = np.zeros(nfull + 1, dtype=bool)
                np.logical_or.at(ga, np.arange(B) // G, active)
                tail = int(ga[nfull]) if B % G else 0
                idx = self.torch.as_tensor(np.flatnonzero(ga[:nfull]).astype(np.int32), device=self.dev)
            if ld:
                if transpose:
                    raise ValueError('row-major vectors are solved without transposition')
                _lib.call('sawfem_mf_solve_groups_rows', self._plan, stream_ptr(), B, R, arena.data_ptr(), G,
                          uarena.data_ptr(), vec.data_ptr(), ld, nb, None if idx is None else idx.data_ptr(),
                          0 if idx is None else int(idx.numel()), tail)
                return
            _lib.call('sawfem_mf_solve_groups', self._plan, stream_ptr(), B, R, arena.data_ptr(), G,
                      uarena.data_ptr(), vec.data_ptr(), transpose, nb, None if idx is None else idx.data_ptr(),
                      0 if idx is None else int(idx.numel()), tail)
            return
        if ld:
            raise ValueError('row-major vectors need grouped factorizations')
        if active is not None:
            idx = self.torch.as_tensor(np.flatnonzero(active).astype(np.int32), device=self.dev)
        _lib.call('sawfem_mf_solve', self._plan, stream_ptr(), prec, B, R, arena.data_ptr(), uarena.data_ptr(),
                  vec.data_ptr(), transpose, nb, None if idx is None else idx.data_ptr(),
                  0 if idx is None else int(idx.numel()))

    def pivot_stats(self, reset=True):
        """(pivots the guard replaced, smallest |pivot| / row scale) since the last reset
        (``sawfem_mf_pivot_stats``)."""
        import ctypes as C
        cnt, rmin = C.c_int64(), C.c_double()
        _lib.call('sawfem_mf_pivot_stats', self._plan, int(reset), C.byref(cnt), C.byref(rmin))
        return int(cnt.value), float(rmin.value)

    def _refine(self, prec, nb, arena, uarena, X, transpose, res, rhst, tslot, measure, ir_tol, max_ir, dwr=None,
                rows=None, update=None):
        """Iterative refinement of X ([B][R][n], holding the first solve) in place; with
        ``rows=(B, R, ld)`` X is row-major [n][ld] (vector b R + r in column b R + r; grouped
        solves only), and so are rhst and the residuals.  ``update(r)``, when given, applies the
        correction X += r and returns what ``measure()`` would then (one fused pass).

        Each step: FP64 residual of every frequency, a solve of the frequencies still active,
        update.  A frequency converges when the relative change of every quantity ``measure()``
        returns is at most ``ir_tol`` (the remaining error is then below ir_tol for any rate
        under 1/2, and below 100 ir_tol even for a slow tail at rate 0.99); from the second step
        on, one whose change above ir_tol does not halve has stalled (left to the complex128
        refactorization).  Converged and stalled frequencies are left out of the solve (their
        correction is zero).

        ``dwr(r)``, when given, returns for every quantity the dual-weighted residual: the
        adjoint solution contracted with the residual r = b - A x of the current iterate, which
        is the quantity's error to first order (for Q = k^T x, Q - Q* = -lambda*^T r exactly, with
        lambda* = A^-T k; the computed lambda replaces lambda* at second-order cost).  A frequency
        whose relative dual-weighted residuals are all within ``ir_tol`` converges on the residual
        alone, before that step's solve.

        Returns (values, last changes, error estimates, steps, history of the largest change of
        the first quantity); ``self._converged`` marks the converged frequencies and
        ``self._dwr_ok`` those accepted on their dual-weighted residual.  The estimates are the
        dual-weighted residuals where accepted on them, :func:`ir_error` otherwise.
        """
        torch = self.torch
        s = stream_ptr()
        if rows:
            B, R, ld = rows
        else:
            B, ld = X.shape[0], 0
            R = X.shape[1] if X.dim() == 3 else 1
        vals = measure()
        dqs = [np.full(B, np.inf) for _ in vals]
        errs = [np.full(B, np.inf) for _ in vals]
        active = np.ones(B, dtype=bool)
        dwr_ok = np.zeros(B, dtype=bool)
        r = torch.zeros_like(X) if ld else torch.empty_like(X)     # row-major: padding columns stay 0
        hist, applied = [], 0
        self._solved = [B]                 # frequencies solved per pass (first solve, then each step)
        self._stalled = np.zeros(B, dtype=bool)
        for steps in range(1, max_ir + 1):
            res(X, rhst, r, tslot, active, R)
            if dwr is not None:
                eta = [np.abs(a) / np.maximum(np.abs(v), 1e-300) for a, v in zip(dwr(r), vals)]
                ok = active & np.all([e <= ir_tol for e in eta], axis=0)
                for k, e in enumerate(eta):
                    errs[k][ok] = e[ok]
                dwr_ok |= ok
                active &= ~ok
                if not active.any():
                    break
            self._solved.append(int(active.sum()))
            sel = None
            if not active.all():
                if ld:
                    cols = (np.flatnonzero(~active)[:, None] * R + np.arange(R)).ravel()
                    r[:, torch.as_tensor(cols, device=self.dev)] = 0
                else:
                    r[torch.as_tensor(~active, device=self.dev)] = 0
                sel = active
            self._solve(prec, B, R, arena, uarena, r, transpose, nb, sel, ld=ld)
            if update is not None:
                new = update(r)                          # X += r and measure() in one pass
            else:
                _lib.call('sawfem_axpy', s, X.numel(), X.data_ptr(), r.data_ptr())
                new = measure()
            applied = steps
            stalled = np.zeros(B, dtype=bool)
            for k in range(len(vals)):
                dq = np.abs(new[k] - vals[k]) / np.maximum(np.abs(new[k]), 1e-300)
                e = ir_error(dq, dqs[k])
                # a correction that does not shrink by half after the first step: refinement with
                # these factors does not converge here (left to the complex128 refactorization)
                if steps >= 2:
                    stalled |= active & (dq > 0.5 * dqs[k]) & (dq > ir_tol)
                errs[k][active] = e[active]
                dqs[k][active] = dq[active]
                vals[k] = new[k]
            hist.append(float(dqs[0][active].max()))
            active = active & np.any([d > ir_tol for d in dqs], axis=0) & ~stalled & ~self._stalled
            self._stalled |= stalled
            if not active.any():
                break
        self._dwr_ok = dwr_ok
        self._converged = dwr_ok | (np.all([d <= ir_tol for d in dqs], axis=0) & ~self._stalled)
        return vals, dqs, errs, applied, hist

    def _solve_single(self, om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir):
        """x = A^-1 rhs refined against the FP64 operator until Q's estimated error <= ir_tol."""
        torch = self.torch
        s = stream_ptr()
        B, n = rhs.shape
        x = rhs.clone()
        self._solve(prec, B, 1, arena, uarena, x, 0, nb)
        Qd = torch.empty(B, dtype=torch.complex128, device=self.dev)

        def measure():
            _lib.call('sawfem_charge', s, n, B, self.kts_a.data_ptr(), self.kts_z.real, self.kts_z.imag, float(V),
                      x.data_ptr(), Qd.data_ptr())
            return [Qd.cpu().numpy().copy()]
        res = self._residual_op(B, om)
        (Q,), (dq,), (err,), steps, hist = self._refine(prec, nb, arena, uarena, x, 0, res, res.to_t(rhs), None,
                                                         measure, ir_tol, max_ir)
        return dict(x=x, Q=Q, ir_steps=steps, dq=dq, err=err, dq_hist=hist, solved=list(self._solved),
                    converged=self._converged.copy())

    def _solve_paired(self, om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir):
        """Forward and adjoint systems as two right-hand sides of one refined solve.

        With S = D A symmetric (see :meth:`scaled_symmetric`), lambda = A^-T kts_a =
        D z where A z = D^-1 kts_a.  Both columns are refined together against the FP64
        operator: Q's estimated error and that of the duality product lambda^T rhs must
        both reach ``ir_tol``.
        """
        torch = self.torch
        s = stream_ptr()
        B, n = rhs.shape
        d = self._dscale
        if self._group > 1:
            return self._solve_paired_rows(om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir)
        X = torch.stack([rhs, (self.kts_a / d).unsqueeze(0).expand(B, n)], dim=1).contiguous()   # [B][2][n]
        rhs2 = X.clone()
        self._solve(prec, B, 2, arena, uarena, X, 0, nb)
        Qd = torch.empty(2 * B, dtype=torch.complex128, device=self.dev)

        w3 = torch.empty((3, B), dtype=torch.complex128, device=self.dev)
        ev = 16 * n                      # bytes per [n] vector

        def wdot(a_ptr, sa, c_ptr, sc, out):
            # out[b] = sum_i a_b[i] d[i] c_b[i] (one pass over both vectors)
            _lib.call('sawfem_wdot', s, n, B, a_ptr, sa, d.data_ptr(), c_ptr, sc, None, out.data_ptr())

        def measure():
            _lib.call('sawfem_charge', s, n, 2 * B, self.kts_a.data_ptr(), self.kts_z.real, self.kts_z.imag, float(V),
                      X.data_ptr(), Qd.data_ptr())
            wdot(X.data_ptr() + ev, 2 * n, rhs.data_ptr(), n, w3[0])              # mu = (D lambda)^T b
            return [Qd[0::2].cpu().numpy().copy(), w3[0].cpu().numpy()]
        def dwr(r):
            # Q - Q* = -lambda^T r_x (forward residual); mu - mu* = -(D x)^T r_z (adjoint residual)
            wdot(X.data_ptr() + ev, 2 * n, r.data_ptr(), 2 * n, w3[1])
            wdot(X.data_ptr(), 2 * n, r.data_ptr() + ev, 2 * n, w3[2])
            e = w3[1:].cpu().numpy()
            return [e[0], e[1]]
        res = self._residual_op(2 * B, om.repeat_interleave(2))
        (Q, mu), (dq, dm), (err, aerr), steps, hist = self._refine(prec, nb, arena, uarena, X, 0, res, res.to_t(rhs2),
                                                                  None, measure, ir_tol, max_ir, dwr=dwr)
        x = X[:, 0].contiguous()
        lam = (X[:, 1] * d).contiguous()
        del X, rhs2
        return self._paired_out(omegas, V, x, lam, Q, steps, dq, err, hist, dm, aerr)

    def _solve_paired_rows(self, om, omegas, V, prec, nb, arena, uarena, rhs, ir_tol, max_ir):
        """:meth:`_solve_paired` with grouped factorizations, its vectors kept row-major for the
        whole refinement: column 2b holds frequency b's forward vector, column 2b + 1 its adjoint
        (X: [n][ld], ld = 2B padded to 64).  The grouped solves then gather whole lines, and the
        residual reads and writes X and r in place (no transposes).  The charge and duality
        products run in the summation order of the [B][2][n] path, so every result is the same
        bit for bit."""
        torch = self.torch
        s = stream_ptr()
        B, n = rhs.shape
        d = self._dscale
        nv = 2 * B
        ld = -(-nv // 64) * 64
        c128 = dict(dtype=torch.complex128, device=self.dev)
        # [B][2][n] right-hand sides transposed (tiled) into the row-major array
        V2 = torch.stack([rhs, (self.kts_a / d).unsqueeze(0).expand(B, n)], dim=1)
        X = torch.zeros((n, ld), **c128)
        _lib.call('sawfem_transpose', s, nv, n, n, ld, V2.data_ptr(), X.data_ptr())
        del V2
        rhs2 = X.clone()
        self._solve(prec, B, 2, arena, uarena, X, 0, nb, ld=ld)
        Qd = torch.empty(nv, **c128)
        w4 = torch.empty((4, B), **c128)
        work = torch.empty(3072 * B, **c128)             # partial sums of the row-major products

        def wdots(c, out1, out2):
            # out1[b] = sum_i X[i][2b + 1] d[i] c[i][2b], out2[b] = sum_i X[i][2b] d[i] c[i][2b + 1]
            _lib.call('sawfem_wdot_pairs_rows', s, n, B, ld, X.data_ptr(), d.data_ptr(), c.data_ptr(), work.data_ptr(),
                      out1.data_ptr(), out2.data_ptr())

        def measure():
            _lib.call('sawfem_charge_rows', s, n, nv, ld, self.kts_a.data_ptr(), self.kts_z.real, self.kts_z.imag,
                      float(V), X.data_ptr(), None, work.data_ptr(), Qd.data_ptr())
            wdots(rhs2, w4[0], w4[3])                                             # mu = (D lambda)^T b
            return [Qd[0::2].cpu().numpy().copy(), w4[0].cpu().numpy()]
        def dwr(r):
            # Q - Q* = -lambda^T r_x (forward residual); mu - mu* = -(D x)^T r_z (adjoint residual)
            wdots(r, w4[1], w4[2])
            e = w4[1:3].cpu().numpy()
            return [e[0], e[1]]
        def update(r):
            # X += r, then measure()'s charges and mu, in one pass over X, r and the right-hand sides
            _lib.call('sawfem_rows_update_measure', s, n, B, ld, X.data_ptr(), r.data_ptr(), self.kts_a.data_ptr(),
                      self.kts_z.real, self.kts_z.imag, float(V), d.data_ptr(), rhs2.data_ptr(), work.data_ptr(),
                      Qd.data_ptr(), w4[0].data_ptr())
            return [Qd[0::2].cpu().numpy().copy(), w4[0].cpu().numpy()]
        res = self._residual_op(nv, om.repeat_interleave(2), ld=ld)
        (Q, mu), (dq, dm), (err, aerr), steps, hist = self._refine(prec, nb, arena, uarena, X, 0, res, rhs2, None,
                                                                  measure, ir_tol, max_ir, dwr=dwr, rows=(B, 2, ld),
                                                                  update=update)
        XT = torch.empty((B, 2, n), **c128)              # back to [B][2][n] by one tiled transpose
        _lib.call('sawfem_transpose', s, n, nv, ld, n, X.data_ptr(), XT.data_ptr())
        x = XT[:, 0].contiguous()
        lam = (XT[:, 1] * d).contiguous()
        del X, XT, rhs2, work
        return self._paired_out(omegas, V, x, lam, Q, steps, dq, err, hist, dm, aerr)

    def _paired_out(self, omegas, V, x, lam, Q, steps, dq, err, hist, dm, aerr):
        """The paired solve's result: the refinement record, dY11/dp from x and lambda ([B][n])."""
        out = dict(x=x, Q=Q, ir_steps=steps, dq=dq, err=err, dq_hist=hist, solved=list(self._solved),
                   converged=self._converged.copy(), dwr_accepted=int(self._dwr_ok.sum()),
                   dwr_mask=self._dwr_ok.copy())
        out.update(self._contract(omegas, V, x, lam))
        out.update(adj_dq=dm, adj_err=aerr, adj_steps=steps, lam=lam, paired=True)
        return out

    def _residual_op(self, B, om, ld=0):
        """r = rhs - A(w) x (or A^T with tslot) through the frequency-contiguous residual kernel
        (row-major operands [n][ld] when ``ld``)."""
        return _ResidualOp(self, B, om, ld)

    def _adjoint(self, om, omegas, V, prec, nb, arena, uarena, rhs, x, ir_tol, max_ir):
        """lambda = A^-T kts_a by transposed solves (refined with the transposed FP64 residual
        until the duality product lambda^T b converges), then K12 and the parameter contraction."""
        s = stream_ptr()
        B, n = x.shape
        gvec = self.kts_a.unsqueeze(0).expand(B, n).contiguous()
        lam = gvec.clone()
        self._solve(prec, B, 1, arena, uarena, lam, 1, nb)
        measure = lambda: [(lam * rhs).sum(dim=1).cpu().numpy()]
        res = self._residual_op(B, om)
        _, (dq,), (err,), steps, _ = self._refine(prec, nb, arena, uarena, lam, 1, res, res.to_t(gvec), self.tslot,
                                                  measure, ir_tol, max_ir)
        out = self._contract(omegas, V, x, lam)
        out.update(adj_dq=dq, adj_err=err, adj_steps=steps, lam=lam, adj_converged=self._converged.copy())
        return out

    def _contract(self, omegas, V, x, lam):
        """dY11/dp from the forward solution x and the adjoint lambda (element sensitivities
        on the GPU, parameter contraction on the host)."""
        torch = self.torch
        g = self._grad
        s = stream_ptr()
        B, n = x.shape
        ndu = self.layout.nd - 1
        NT = (2 * ndu) ** 2 + 4 * ndu + 4
        work = torch.empty(max(1, B * g['ne0'] * NT), dtype=torch.complex128, device=self.dev)
        W = torch.empty(B * NT, dtype=torch.complex128, device=self.dev)
        tags, kinds, par = g['pml']
        from .piezo_fem import SCALE
        _lib.call('sawfem_adjoint_elem', s, g['ne0'], g['elems'].data_ptr(), g['conn'].data_ptr(), g['xy'].data_ptr(),
                  g['tag'].data_ptr(), g['npml'], tags.ctypes.data, kinds.ctypes.data, par.ctypes.data,
                  g['omega_top'], ndu, g['N'], g['pos_rep'].data_ptr(), g['zmask'].data_ptr(), float(V), SCALE, B, n,
                  x.data_ptr(), lam.data_ptr(), work.data_ptr(), W.data_ptr())
        Wh = W.cpu().numpy().reshape(B, NT)
        return dict(dY=self._param_gradient(Wh, omegas, V))

    def auto_batch(self, precision='mixed', gradient=False, budget=None, share=None):
        """Frequencies per chunk (:meth:`chunk_bytes`) that fit ``budget`` bytes, by default the
        device memory available above the reserve (:func:`hbm_available`, counting the current
        workspace as free)."""
        if budget is None:
            ws = getattr(self, '_ws', None)
            budget = hbm_available(0 if ws is None else ws.numel())
        share = max(1, int(self.share if share is None else share)) if precision == 'mixed' else 1
        per = self.chunk_bytes(1, precision, gradient, share)
        if budget < per:
            raise ValueError(f'one frequency needs {per / 1e9:.1f} GB of device memory; '
                             f'{max(budget, 0) / 1e9:.1f} GB are available above the reserve')
        # chunk_bytes(hi) > budget: its arena alone exceeds the budget
        lo, hi = 1, share * (max(1, int(budget // self.arena_bytes(1, precision))) + 1)
        while hi - lo > 1:                      # chunk_bytes grows with the batch: bisect
            mid = (lo + hi) // 2
            lo, hi = (mid, hi) if self.chunk_bytes(mid, precision, gradient, share) <= budget else (lo, mid)
        return lo

    def sweep_charges(self, fre, V=1.0, precision='mixed', batch=None, gradient=False, on_chunk=None, **kw):
        """Charges Q(f) for all frequencies, in chunks of ``batch`` frequencies (default: as many
        as fit the free HBM, see :meth:`auto_batch`).

        With ``gradient=True`` returns (Q, dY, info) where dY (P, n_params) is dY11/dp.
        ``on_chunk(done, total)`` is called after each chunk (progress and cancellation points).
        """
        fre = np.asarray(fre, dtype=np.float64).ravel()
        if fre.size == 0:                      # nothing to solve: empty results, no launch
            Q = np.empty(0, dtype=complex)
            if gradient:
                return Q, np.empty((0, self.n_params), dtype=complex), []
            return Q, []
        if not batch:
            nb = self.auto_batch(precision, gradient)
            batch = -(-fre.size // -(-fre.size // nb))          # equal chunks
        Q = np.empty(fre.size, dtype=complex)
        dY = np.empty((fre.size, self.n_params), dtype=complex) if gradient else None
        info = []
        for i0 in range(0, fre.size, batch):
            res = self.solve_batch(2 * np.pi * fre[i0:i0 + batch], V, precision=precision, gradient=gradient, **kw)
            Q[i0:i0 + batch] = res['Q']
            item = dict(ir_steps=res['ir_steps'], dq_max=float(np.max(res['dq'])),
                        err_est_max=float(np.max(res['err'])), refactored=int(res['refactored'].sum()),
                        perturbed=res['perturbed'], pivot_min_ratio=res['pivot_min_ratio'], share=res['share'],
                        factorizations=res['factorizations'], own_factors=int(res['own_factors'].sum()))
            if gradient:
                dY[i0:i0 + batch] = res['dY']
                item['adj_dq_max'] = float(np.max(res['adj_dq']))
                item['adj_err_est_max'] = float(np.max(res['adj_err']))
                item['adj_steps'] = res['adj_steps']
                item['paired_adjoint'] = bool(res.get('paired', False))
            item['points'] = int(min(batch, fre.size - i0))
            item['dq_hist'] = res.get('dq_hist')
            item['solved_per_pass'] = res.get('solved')
            if 'dwr_accepted' in res:
                item['accepted_on_residual'] = res['dwr_accepted']
            info.append(item)
            del res
            if on_chunk is not None:
                on_chunk(min(i0 + batch, fre.size), fre.size)
        if gradient:
            return Q, dY, info
        return Q, info

    def expand(self, x, V=1.0):
        return self.bloch.expand(np.asarray(x), V)

    def solve(self, omega: float, V: float = 1.0):
        """(disp, Q, |iωQ/V|) for one frequency (``sweep.py:335-364``), FP64 path."""
        res = self.solve_batch([omega], V, precision='fp64')
        disp = self.expand(res['x'][0].cpu().numpy(), V)
        Q = complex(res['Q'][0])
        return disp, Q, abs(1j * omega * Q / V)

    def sweep(self, fre, V: float = 1.0, n_jobs: int = 1, progress: bool = False, precision='mixed', batch=None):
        """(Q, |Y|) over frequencies ``fre`` in Hz (``sweep.py:366-393``); all points batched."""
        Q, _ = self.sweep_charges(fre, V, precision=precision, batch=batch)
        return Q, np.abs(1j * 2 * np.pi * np.asarray(fre) * Q / V)
