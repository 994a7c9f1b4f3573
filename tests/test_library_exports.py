"""The C-ABI library loads and exports every symbol declared in include/sawfem.h (CPU)."""
import ctypes
import re
from pathlib import Path

HEADER = Path(__file__).resolve().parents[1] / 'include' / 'sawfem.h'


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'\b(sawfem_[a-z0-9_]+)\s*\(', text)))


def test_every_declared_symbol_is_exported(built_lib):
    lib = ctypes.CDLL(str(built_lib))
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    from sawsim_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared())


def test_invalid_arguments_fail_loudly(built_lib):
    """Argument checks run on the host before any GPU work."""
    import pytest
    from sawsim_amd import _lib
    with pytest.raises(_lib.SawfemError, match='invalid'):
        _lib.call('sawfem_elem_q9', None, 1, 7, None, None, None, None, None, None, None, None, 0, None, None,
                  None, 1.0, 1.0, None, None)
    with pytest.raises(_lib.SawfemError):
        _lib.call('sawfem_mf_factor', None, None, 0, 0, None, 0.0, None, None, None, None, 32)
