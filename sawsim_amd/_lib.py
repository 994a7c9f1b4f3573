"""ctypes binding of ``libsawfem_hip.so`` (the C ABI declared in ``include/sawfem.h``).

The product path has no CPU fallback: every function here raises if the HIP library
is missing or a call returns a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_NAME = 'libsawfem_hip.so'
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    'sawfem_last_error': [],
    'sawfem_elem_q9': [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _d, _d, _vp, _vp],
    'sawfem_gather_kmd': [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'sawfem_elem_h27': [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _d, _d, _vp, _vp],
    'sawfem_gather_h27': [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'sawfem_adjoint_h27': [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _d, _i, _vp, _vp, _d, _d, _i, _i, _vp, _vp,
                           _vp, _vp],
    'sawfem_gather3': [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'sawfem_mf_create': [C.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _i64, _vp,
                         _vp, _vp, _vp],
    'sawfem_mf_destroy': [_vp],
    'sawfem_analyse': [C.POINTER(_vp), _i64, _vp, _vp, _vp, _vp, _i64, _vp, _d, _i64],
    'sawfem_bloch_layout': [C.POINTER(_vp), _i, _i64, _i, _vp, _vp, _vp, _vp, _i64, _vp, _vp],
    'sawfem_transpose_slots': [_i64, _vp, _vp, _vp],
    'sawfem_host_size': [_vp, _i],
    'sawfem_host_copy': [_vp, _i, _vp],
    'sawfem_host_destroy': [_vp],
    'sawfem_mf_create_sym': [C.POINTER(_vp), _vp],
    'sawfem_mf_arena_elems': [_vp],
    'sawfem_mf_uarena_elems': [_vp],
    'sawfem_mf_perturbed': [_vp, _i],
    'sawfem_mf_pivot_stats': [_vp, _i, _vp, _vp],
    'sawfem_mf_set_small': [_vp, _i, _i],
    'sawfem_mf_set_solve_full': [_vp, _i],
    'sawfem_mf_set_panel_lu': [_vp, _i],
    'sawfem_mf_set_tri_lds': [_vp, _i],
    'sawfem_mf_profile_levels': [_vp, _vp, _i],
    'sawfem_mf_factor': [_vp, _vp, _i, _i, _vp, _d, _vp, _vp, _vp, _vp, _i],
    'sawfem_mf_factor_graph': [_vp, _vp, _i, _i, _vp, _d, _vp, _vp, _vp, _vp, _i],
    'sawfem_mf_graph_nodes': [_vp],
    'sawfem_mf_solve': [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _i],
    'sawfem_mf_solve_shared': [_vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _vp, _i],
    'sawfem_mf_solve_groups': [_vp, _vp, _i, _i, _vp, _i, _vp, _vp, _i, _i, _vp, _i, _i],
    'sawfem_mf_solve_groups_rows': [_vp, _vp, _i, _i, _vp, _i, _vp, _vp, _i64, _i, _vp, _i, _i],
    'sawfem_mf_set_group_rc': [_vp, _i],
    'sawfem_mf_set_group_tri': [_vp, _i, _i],
    'sawfem_mf_set_group_big': [_vp, _i],
    'sawfem_mf_profile': [_vp, _i],
    'sawfem_mf_profile_read': [_vp, _vp, _vp],
    'sawfem_rhs': [_vp, _i, _i, _vp, _d, _d, _vp, _vp, _vp, _vp],
    'sawfem_residual': [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp],
    'sawfem_residual_t': [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp],
    'sawfem_transpose_map': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp],
    'sawfem_mf_profile_grp': [_vp, _vp, _vp, _vp],
    'sawfem_wdot': [_vp, _i, _i, _vp, _i64, _vp, _vp, _i64, _vp, _vp],
    'sawfem_residual_sv': [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp],
    'sawfem_charge_rows': [_vp, _i, _i, _i64, _vp, _d, _d, _d, _vp, _vp, _vp, _vp],
    'sawfem_wdot_rows': [_vp, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'sawfem_wdot_pairs_rows': [_vp, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    'sawfem_rows_update_measure': [_vp, _i, _i, _i64, _vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, _vp, _vp],
    'sawfem_transpose': [_vp, _i, _i, _i, _i, _vp, _vp],
    'sawfem_charge': [_vp, _i, _i, _vp, _d, _d, _d, _vp, _vp],
    'sawfem_axpy': [_vp, _i64, _vp, _vp],
    'sawfem_adjoint_elem': [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _d, _i, _i, _vp, _vp, _d, _d, _i, _i,
                            _vp, _vp, _vp, _vp],
    'sawfem_sweep_chunk_bytes': [_vp, _i, _i, _i],
    'sawfem_sweep_chunk': [_vp, _vp, _vp, _i, _vp, _d, _i, _d, _i, _i, _vp, _i64, _vp, _vp, _vp],
    'sawfem_sweep_chunk_share_bytes': [_vp, _i, _i, _i, _i],
    'sawfem_sweep_chunk_share': [_vp, _vp, _vp, _i, _vp, _d, _i, _i, _d, _i, _i, _vp, _i64, _vp, _vp, _vp],
    'sawfem_adjoint_elem_noalloc': [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _d, _i, _i, _vp, _vp, _d, _d, _i,
                                    _i, _vp, _vp, _vp],
}
_RESTYPE = {'sawfem_last_error': C.c_char_p, 'sawfem_mf_arena_elems': _i64, 'sawfem_mf_uarena_elems': _i64,
            'sawfem_sweep_chunk_bytes': _i64, 'sawfem_sweep_chunk_share_bytes': _i64, 'sawfem_host_size': _i64, 'sawfem_mf_graph_nodes': _i64}

# status codes of the C ABI (include/sawfem.h)
OK, INVALID, PLAN_LIMIT, HIP_ERROR, OUT_OF_MEMORY, NON_FINITE, REFACTORED = 0, 1, 2, 3, 4, 5, 6


class SweepOperator(C.Structure):
    """``sawfem_operator``: the active operator of one cell (device pointers)."""
    _fields_ = [('n', _i), ('indptr', _vp), ('indices', _vp), ('Kaa', _vp), ('Maa', _vp), ('Daa', _vp),
                ('rK', _vp), ('rM', _vp), ('rD', _vp), ('kts_a', _vp), ('kts_z_re', _d), ('kts_z_im', _d),
                ('beta', _d)]

_lib = None


class SawfemError(RuntimeError):
    pass


def load():
    """Load the HIP library (once); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get('SAWFEM_LIB', LIB_PATH))
    if not path.exists():
        raise SawfemError(f'{path} is missing: build it with `python -c "import __graft_entry__ as g; g.build()"`')
    lib = C.CDLL(str(path))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, _i)
    _lib = lib
    return lib


def call(name, *args):
    """Call a status-returning entry point; raise with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.sawfem_last_error().decode(errors='replace')
        raise SawfemError(f'{name} failed with status {rc}: {msg}')
    return rc


def call_value(name, *args):
    """Call an entry point that returns a value (negative on failure); raise on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc < 0:
        msg = lib.sawfem_last_error().decode(errors='replace')
        raise SawfemError(f'{name} failed: {msg}')
    return rc


def raw(name, *args):
    return getattr(load(), name)(*args)


def host_arrays(handle, spec):
    """Copy the arrays ``spec`` = {name: (id, dtype)} out of a host result and free it."""
    try:
        out = {}
        for name, (what, dt) in spec.items():
            a = np.empty(call_value('sawfem_host_size', handle, what), dtype=dt)
            call('sawfem_host_copy', handle, what, a.ctypes.data)
            out[name] = a
        return out
    finally:
        call('sawfem_host_destroy', handle)
