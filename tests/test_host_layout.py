"""Host integer work of the product path against the oracle (CPU)."""
import numpy as np
import pytest

import helpers
import mf_emulator
from oracle import sweep as osw
from sawsim_amd import materials as mt, workloads
from sawsim_amd.layout import BlochLayout, KLayout
from sawsim_amd.piezo_fem import Model
from sawsim_amd.plan import kernel_work
from sawsim_amd.symbolic import analyse
from sawsim_amd.sweep import bc_periodic


def _cells():
    return [workloads.double_layer(), workloads.config1(mesh_um=0.12, points=3)]


@pytest.mark.parametrize('cell', _cells(), ids=['double_layer', 'config1_coarse'])
def test_structural_pattern_and_bloch_maps_match_oracle(cell):
    m = cell.model
    nd = cell.ndisp + 1
    kl = KLayout(m.quads, m.N, nd)
    K, M = helpers.oracle_assemble(cell, 2 * np.pi * cell.band_hz[1])
    indptr, indices = kl.csr()
    # the reference's COO -> CSR keeps every element slot: identical structural pattern
    assert np.array_equal(indptr, K.indptr) and np.array_equal(indices, K.indices)
    bcs = bc_periodic(m, cell.mesh.lines, ndisp=cell.ndisp)
    obcs = osw.bc_periodic(m.pos_m, cell.mesh.lines, m.N, ndisp=cell.ndisp)
    for a, b in zip(bcs[:3], obcs[:3]):
        assert np.array_equal(a, b)
    assert all(np.array_equal(g, h) for g, h in zip(bcs[3], obcs[3]))
    bl = BlochLayout(kl, *bcs)
    eng = osw.FrequencySolver(K, M, *obcs)
    assert np.array_equal(bl.active, eng.active)
    o = np.argsort(eng.slaves)
    assert np.array_equal(bl.slaves, eng.slaves[o]) and np.array_equal(bl.masters, eng.masters[o])
    # oracle Km_aa (SpGEMM drops exact zeros) is contained in the structural active pattern
    r = np.repeat(np.arange(bl.n_active), np.diff(bl.aa_indptr))
    ours = set(zip(r.tolist(), bl.aa_indices.tolist()))
    ref = eng.Km_aa.tocoo()
    assert set(zip(ref.row.tolist(), ref.col.tolist())) <= ours
    # gathers reproduce Km_aa, Km@s, kts on the host
    Kv = K.data
    Kaa = np.array([Kv[bl.aa.idx[bl.aa.ptr[j]:bl.aa.ptr[j + 1]]].sum() for j in range(bl.nnz_aa)])
    assert np.abs(Kaa - np.asarray(eng.Km_aa[r, bl.aa_indices]).ravel()).max() <= 1e-12 * np.abs(Kaa).max()
    rK = np.array([Kv[bl.rhs.idx[bl.rhs.ptr[j]:bl.rhs.ptr[j + 1]]].sum() for j in range(bl.n_active)])
    assert np.allclose(rK, eng.Km_s[eng.active], rtol=1e-12, atol=1e-12 * np.abs(rK).max())
    kt = np.array([Kv[bl.kts.idx[bl.kts.ptr[j]:bl.kts.ptr[j + 1]]].sum() for j in range(bl.n_active + 1)])
    x = np.random.default_rng(3).standard_normal(bl.n_active) + 1j
    assert np.isclose(kt[:-1] @ x + kt[-1], eng.kts @ eng.expand(x), rtol=1e-12)


def test_product_materials_match_oracle():
    from oracle import materials as om
    k = workloads.jittered_constants(5)
    m = mt.trigonal_3m(k, 7450.0)
    C, e, eps = om.crystal_tensors('trigonal_3m', k)
    assert np.allclose(m.C, C, rtol=1e-14) and np.allclose(m.e, e) and np.allclose(m.eps, eps)
    for euler in [(0, -42, 0), (10, -38, 44)]:
        got = mt.rotate(euler, m.C, m.e, m.eps)
        ref = om.rotate(euler, C, e, eps)
        for g, r in zip(got, ref):
            assert np.allclose(g, r, rtol=1e-12, atol=1e-12 * np.abs(r).max())
    iso = mt.isotropic(130e9, 0.28, 2330.0)
    assert np.allclose(iso.C, om.isotropic_C(130e9, 0.28), rtol=1e-14)


@pytest.mark.parametrize('nb', [8, 32])
def test_plan_arrays_drive_a_correct_factorization(nb):
    """The exact data flow of the HIP kernels, emulated in NumPy, solves the system."""
    import scipy.sparse.linalg as spla
    cell = workloads.double_layer()
    m = cell.model
    wtop = 2 * np.pi * cell.band_hz[1]
    eng = helpers.oracle_engine(cell, wtop)
    kl = KLayout(m.quads, m.N, 4)
    bl = BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=3))
    sym = analyse(bl.aa_indptr, bl.aa_indices, bl.node_of, bl.comp_of, m.pos_m, period=2.17e-6, leaf_weight=60)
    mf = sym.mf
    r = np.repeat(np.arange(bl.n_active), np.diff(bl.aa_indptr))
    Kaa = np.asarray(eng.Km_aa[r, bl.aa_indices]).ravel()
    Maa = np.asarray(eng.Mm_aa[r, bl.aa_indices]).ravel()
    om_ = 2 * np.pi * np.array([1.7e9, 2.4e9])
    arena = mf_emulator.factor(mf, 2, om_, 0.0, Kaa, Maa, None, nb=nb)
    rhs = np.stack([eng.system(w)[1] for w in om_])
    x = mf_emulator.solve(mf, 2, arena, rhs.copy(), nb=nb)
    # diagonal blocks are applied as explicit triangular inverses (no pivoting): the residual
    # grows with their condition number, ~1e-10 for 32-wide blocks here; a wrong index or map
    # gives an O(1) residual
    tol = 1e-12 if nb <= 8 else 1e-9
    for b, w in enumerate(om_):
        A, _ = eng.system(w)
        assert np.linalg.norm(A @ x[b] - rhs[b]) <= tol * np.linalg.norm(rhs[b])
    work = kernel_work(mf, nb)
    assert 0 < work['gemm_flops'] <= work['factor_flops']
    assert np.isclose(work['factor_flops'], sym.stats['flops_factor'])
    # Schur bytes, front by front: L21 (q x p) and U12 (p x q) read, the q x q block read and written
    sb = sum(8 * (2 * f.p * (f.m - f.p) + 2 * (f.m - f.p) ** 2) for f in sym.fronts if f.m > f.p and f.p > 0)
    assert np.isclose(work['schur_bytes'], sb)


def test_symbolic_covers_every_unknown_once():
    cell = workloads.config1(mesh_um=0.1, points=3)
    m = cell.model
    kl = KLayout(m.quads, m.N, 3)
    bl = BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=2))
    sym = analyse(bl.aa_indptr, bl.aa_indices, bl.node_of, bl.comp_of, m.pos_m, period=2.17e-6)
    assert np.array_equal(np.sort(sym.perm), np.arange(bl.n_active))
    for f in sym.fronts:
        assert f.rows.size == f.m and np.unique(f.rows).size == f.m
        if f.parent >= 0:
            par = sym.fronts[f.parent]
            assert set(f.rows[f.p:].tolist()) <= set(par.rows.tolist())


def _planted_zero_pivot_system():
    """A small cell's active system at one frequency with the first pivot of a leaf front made
    exactly zero (K_gg = w^2 M_gg): LU without pivoting breaks down there although A is
    regular."""
    cell = workloads.double_layer()
    m = cell.model
    wtop = 2 * np.pi * cell.band_hz[1]
    eng = helpers.oracle_engine(cell, wtop)
    kl = KLayout(m.quads, m.N, 4)
    bl = BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=3))
    sym = analyse(bl.aa_indptr, bl.aa_indices, bl.node_of, bl.comp_of, m.pos_m, period=2.17e-6, leaf_weight=60)
    mf = sym.mf
    r = np.repeat(np.arange(bl.n_active), np.diff(bl.aa_indptr))
    Kaa = np.asarray(eng.Km_aa[r, bl.aa_indices]).ravel()
    Maa = np.asarray(eng.Mm_aa[r, bl.aa_indices]).ravel()
    w = 2 * np.pi * 2.0e9
    leaf = int(mf.lvl_fronts[0])
    g = int(mf.rows[mf.rows_ptr[leaf]])          # first pivot row of a childless front
    dslot = int(np.flatnonzero((r == g) & (bl.aa_indices == g))[0])
    Kaa = Kaa.copy()
    Kaa[dslot] = w * w * Maa[dslot]
    assert Kaa[dslot] - w * w * Maa[dslot] == 0
    return mf, bl, r, Kaa, Maa, w


def test_pivot_guard_replaces_a_planted_zero_pivot_and_refinement_recovers():
    """The kernels' pivot guard on the emulated data flow: a planted exact-zero pivot is
    replaced (sqrt(u) times its row scale) and counted and the factors stay finite in both
    precisions; with the perturbed complex128 factors refinement reaches the direct solution of
    the unperturbed system (the complex64 factors go to the complex128 refactorization when
    their refinement stalls); without the guard the same factorization divides by zero."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    mf, bl, r, Kaa, Maa, w = _planted_zero_pivot_system()
    n = bl.n_active
    A = sp.csr_matrix((Kaa - w * w * Maa, bl.aa_indices, bl.aa_indptr), shape=(n, n))
    rhs = np.random.default_rng(0).standard_normal(n) * np.abs(A).max(axis=1).toarray().ravel()
    lu = spla.splu(A.tocsc())
    x_ref = lu.solve(rhs.astype(complex))
    for _ in range(4):     # SuperLU alone is ~1e-5 off on the potential rows of this system
        x_ref = x_ref + lu.solve(rhs - A @ x_ref)
    for dtype in (np.complex128, np.complex64):
        st = {}
        arena = mf_emulator.factor(mf, 1, np.array([w]), 0.0, Kaa, Maa, None, nb=32, dtype=dtype, stats=st)
        assert st['perturbed'] >= 1
        assert np.all(np.isfinite(arena))
        if dtype == np.complex64:
            continue
        x = mf_emulator.solve(mf, 1, arena, rhs[None].astype(complex), nb=32)[0]
        for _ in range(30):
            d = mf_emulator.solve(mf, 1, arena, (rhs - A @ x)[None].astype(complex), nb=32)[0]
            x = x + d
            if np.linalg.norm(d) <= 1e-13 * np.linalg.norm(x):
                break
        assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    # the guard is what keeps the factors finite: with eps = 0 nothing is replaced
    import mf_emulator as em
    keep = em.guard_pivot
    try:
        em.guard_pivot = lambda piv, scale, eps: (piv, False)
        with np.errstate(all='ignore'):
            try:
                bad = em.factor(mf, 1, np.array([w]), 0.0, Kaa, Maa, None, nb=32)
                broke = not np.all(np.isfinite(bad))
            except np.linalg.LinAlgError:       # the zero pivot's triangle is singular
                broke = True
        assert broke
    finally:
        em.guard_pivot = keep


def test_pivot_guard_leaves_regular_pivots_alone():
    """On the unmodified system every pivot clears the guard by orders of magnitude: nothing is
    replaced in either precision."""
    mf, bl, r, Kaa, Maa, w = _planted_zero_pivot_system()
    cell = workloads.double_layer()
    eng = helpers.oracle_engine(cell, 2 * np.pi * cell.band_hz[1])
    Kaa = np.asarray(eng.Km_aa[r, bl.aa_indices]).ravel()
    for dtype in (np.complex128, np.complex64):
        st = {}
        mf_emulator.factor(mf, 1, np.array([w]), 0.0, Kaa, Maa, None, nb=32, dtype=dtype, stats=st)
        assert st['perturbed'] == 0


@pytest.mark.parametrize('cell', [workloads.config1(mesh_um=0.12, points=3), workloads.config2(mesh_um=0.3, points=3)],
                         ids=['config1_coarse', 'config2_coarse'])
def test_residual_groups_share_column_lists(cell):
    """Row groups of the grouped FP64 residual: every active row exactly once, at most four rows
    a group, one merged node a group, identical column lists (checked row by row here), and
    almost every node in one group (the merge keeps a node's unknowns on one pattern)."""
    from sawsim_amd.layout import residual_groups
    m = cell.model
    kl = KLayout(m.quads, m.N, cell.ndisp + 1)
    bl = BlochLayout(kl, *bc_periodic(m, cell.mesh.lines, ndisp=cell.ndisp))
    gptr, grows = residual_groups(bl.aa_indptr, bl.aa_indices, bl.node_of)
    assert np.array_equal(np.sort(grows), np.arange(bl.n_active))
    sizes = np.diff(gptr)
    assert sizes.min() >= 1 and sizes.max() <= 4
    ip, ix = bl.aa_indptr, bl.aa_indices
    for g in range(sizes.size):
        r = grows[gptr[g]:gptr[g + 1]]
        assert np.unique(bl.node_of[r]).size == 1
        for x in r[1:]:
            assert np.array_equal(ix[ip[x]:ip[x + 1]], ix[ip[r[0]]:ip[r[0] + 1]])
    nodes = np.unique(bl.node_of).size
    assert sizes.size <= 1.05 * nodes, (sizes.size, nodes)
