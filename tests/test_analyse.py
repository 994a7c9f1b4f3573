"""The native symbolic analysis (``sawfem_analyse``, host C++) against its NumPy statement
(``tests/symbolic_spec.py``): every ordering, front and plan array bit for bit, on periodic Q9
cells of both unknown layouts, a non-periodic cut, a Hex27 slab (column supervariables), and its
error handling.  No GPU call."""
import ctypes as C

import numpy as np
import pytest

import symbolic_spec as spec
from sawsim_amd import _lib, workloads
from sawsim_amd.layout import BlochLayout, KLayout
from sawsim_amd.symbolic import analyse
from sawsim_amd.sweep import bc_periodic

MF_KEYS = ('m', 'p', 'rows_ptr', 'rows', 'child_ptr', 'child_ids', 'inv_ptr', 'inv', 'lvl_ptr', 'lvl_fronts',
           'ent_front', 'ent_slot', 'ent_loc', 'ent_lvl_ptr')


def _q9_case(cell, nd):
    """Analysis inputs of a plane cell and its Bloch period (as FrequencySolver derives it)."""
    m = cell.model
    kl = KLayout(m.quads, m.N, nd)
    bcs = bc_periodic(m, cell.mesh.lines, ndisp=nd - 1)
    bl = BlochLayout(kl, *bcs)
    xy = np.asarray(m.pos_m)
    g = np.asarray(bcs[3][0]).ravel()
    period = float(abs(xy[g[1] % m.N, 0] - xy[g[0] % m.N, 0]))
    return (bl.aa_indptr, bl.aa_indices, bl.node_of, bl.comp_of, xy), period


def _hex_case():
    """The single-layer Hex27 slab as FrequencySolver3D orders it: (x, z) columns as
    supervariables, 4 interleaved unknowns per node."""
    from sawsim_amd import hex27
    mesh = workloads.hex_template('single').mesh
    kl = KLayout(mesh.hexas, mesh.N, 4, numbering='interleaved')
    bnd = hex27.boundary_idx_vect_3ds(mesh.sets, mesh.pos_m, 3 * mesh.N)
    bl = BlochLayout(kl, bnd.machine['fix'], bnd.electric['zheng'], bnd.electric['fu'],
                     rep=hex27.fold_map(bnd, kl.gdof))
    L, R = bnd.machine['period_L'], bnd.machine['period_R']
    period = float(abs(mesh.pos_m[R[0] // 3, 0] - mesh.pos_m[L[0] // 3, 0]))
    col, col_xz = hex27._columns(mesh.pos_m, period)
    return (bl.aa_indptr, bl.aa_indices, col[bl.node_of], bl.comp_of, col_xz), period


def _compare(args, period, leaf):
    ip, ix, sv, cp, xy = args
    ref = spec.analyse(ip, ix, sv, cp, xy, period=period, leaf_weight=leaf)
    want = spec.build_mf_arrays(ref, ip, ix)
    got = analyse(ip, ix, sv, cp, xy, period=period, leaf_weight=leaf)
    for k in MF_KEYS:
        a, b = getattr(got.mf, k), np.asarray(want[k])
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    assert np.array_equal(got.perm, ref.perm)
    assert np.array_equal(got.order_sv, np.unique(sv)[ref.order_sv])
    assert len(got.levels) == len(ref.levels) and all(np.array_equal(a, b) for a, b in zip(got.levels, ref.levels))
    assert got.stats['root'] == ref.stats['root']
    for f, g in zip(got.fronts, ref.fronts):
        assert f.children == g.children and f.parent == g.parent and f.p == g.p and f.m == g.m
        assert f.level == g.level and np.array_equal(f.rows, g.rows)
    # closed-form factor flops against the front-by-front sum
    ff = sum(float(((g.m - np.arange(g.p) - 1.0) ** 2).sum()) for g in ref.fronts)
    assert np.isclose(got.stats['flops_factor'], 8 * ff, rtol=1e-13)
    return got


@pytest.mark.parametrize('leaf', [60, 96])
def test_native_analysis_matches_spec_double_layer(leaf):
    """ME1 layout ([ux|uy|uz|phi], 4 unknowns per node), periodic cuts."""
    args, period = _q9_case(workloads.double_layer(), 4)
    got = _compare(args, period, leaf)
    assert got.stats['n_fronts'] > 20


def test_native_analysis_matches_spec_config2():
    """ME0 layout ([ux|uy|phi]) of the benchmark's single-layer cell at its full size
    (16,848 elements, 202,144 active unknowns), leaf weight 64 as the engine uses."""
    args, period = _q9_case(workloads.config2(), 3)
    got = _compare(args, period, 64)
    assert got.n == 202144


def test_native_analysis_matches_spec_not_periodic():
    """The same pattern cut without periodicity (x at its weighted median)."""
    args, _ = _q9_case(workloads.double_layer(), 4)
    _compare(args, None, 60)


def test_native_analysis_matches_spec_hex27_slab():
    args, period = _hex_case()
    _compare(args, period, 64)


def test_native_analysis_rejects_bad_input():
    (ip, ix, sv, cp, xy), _ = _q9_case(workloads.double_layer(), 4)
    with pytest.raises(_lib.SawfemError, match='supervariable id'):
        analyse(ip, ix, sv + xy.shape[0], cp, xy, period=2.17e-6)
    bad = ix.copy()
    bad[3] = sv.size + 5
    with pytest.raises(_lib.SawfemError, match='column index'):
        analyse(ip, bad, sv, cp, xy, period=2.17e-6)
    with pytest.raises(ValueError):
        analyse(ip[:-1], ix, sv, cp, xy)
    h = C.c_void_p()
    assert _lib.raw('sawfem_analyse', C.byref(h), 0, None, None, None, None, 0, None, 0.0, 0) == _lib.INVALID
    assert _lib.raw('sawfem_host_size', None, 0) == -1


def _bloch_check(kl, bl):
    is_z = np.zeros(kl.gdof, dtype=bool)
    is_z[bl.zheng] = True
    is_f = np.zeros(kl.gdof, dtype=bool)
    is_f[bl.fu] = True
    want = spec.bloch_arrays(kl, bl.rep, bl.pos, bl.n_active, is_z, is_f)
    got = dict(aa_indptr=bl.aa_indptr, aa_indices=bl.aa_indices, aa_ptr=bl.aa.ptr, aa_idx=bl.aa.idx,
               rhs_ptr=bl.rhs.ptr, rhs_idx=bl.rhs.idx, kts_ptr=bl.kts.ptr, kts_idx=bl.kts.idx)
    for k, v in want.items():
        assert got[k].dtype == np.int64 and np.array_equal(got[k], v), k


@pytest.mark.parametrize('which', ['double_layer', 'config2'])
def test_native_bloch_layout_matches_spec_q9(which):
    """Component numbering, periodic merge groups: the pattern and the three gather lists."""
    cell = getattr(workloads, which)()
    m = cell.model
    kl = KLayout(m.quads, m.N, cell.ndisp + 1)
    _bloch_check(kl, BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=cell.ndisp)))


def test_native_bloch_layout_matches_spec_hex27():
    """Interleaved numbering and the two-stage fold given as a representative map."""
    from sawsim_amd import hex27
    mesh = workloads.hex_template('single').mesh
    kl = KLayout(mesh.hexas, mesh.N, 4, numbering='interleaved')
    bnd = hex27.boundary_idx_vect_3ds(mesh.sets, mesh.pos_m, 3 * mesh.N)
    _bloch_check(kl, BlochLayout(kl, bnd.machine['fix'], bnd.electric['zheng'], bnd.electric['fu'],
                                 rep=hex27.fold_map(bnd, kl.gdof)))


def test_native_bloch_layout_rejects_bad_input():
    cell = workloads.double_layer()
    m = cell.model
    kl = KLayout(m.quads, m.N, 4)
    bl = BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=3))
    h = C.c_void_p()
    z = np.zeros(kl.gdof, dtype=np.uint8)
    pos = bl.pos.copy()
    pos[0] = bl.n_active                              # beyond the active count
    rc = _lib.raw('sawfem_bloch_layout', C.byref(h), 4, kl.N, 0, kl.node_ptr.ctypes.data, kl.node_idx.ctypes.data,
                  bl.rep.ctypes.data, pos.ctypes.data, bl.n_active, z.ctypes.data, z.ctypes.data)
    assert rc == _lib.INVALID and b'out of range' in _lib.load().sawfem_last_error()
    assert _lib.raw('sawfem_bloch_layout', C.byref(h), 4, kl.N, 1, None, None, None, None, 0, None, None) == 1


def test_transpose_slots_matches_sorted_search():
    """sawfem_transpose_slots against a search of the sorted (row, column) keys, and its
    rejection of a pattern that is not structurally symmetric."""
    (ip, ix, _, _, _), _ = _q9_case(workloads.double_layer(), 4)
    n = ip.size - 1
    r = np.repeat(np.arange(n, dtype=np.int64), np.diff(ip))
    key, tkey = r * n + ix, ix * n + r
    want = np.searchsorted(key, tkey)
    assert np.array_equal(key[want], tkey)
    got = np.empty(ix.size, dtype=np.int64)
    _lib.call('sawfem_transpose_slots', n, ip.ctypes.data, ix.ctypes.data, got.ctypes.data)
    assert np.array_equal(got, want)
    # drop one off-diagonal entry: its transpose has no partner
    e = int(np.flatnonzero(r != ix)[5])
    ip2 = ip.copy()
    ip2[r[e] + 1:] -= 1
    ix2 = np.delete(ix, e)
    out = np.empty(ix2.size, dtype=np.int64)
    assert _lib.raw('sawfem_transpose_slots', n, ip2.ctypes.data, ix2.ctypes.data, out.ctypes.data) == _lib.INVALID
