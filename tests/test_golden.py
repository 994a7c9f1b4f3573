"""Oracle and host maps against the golden vectors written from the reference (CPU)."""
import numpy as np
import scipy.sparse as sp

import helpers
from golden_cells import tiny_cell
from sawsim_amd.layout import BlochLayout, KLayout
from sawsim_amd.sweep import bc_periodic


def _csr(g, name, n):
    return sp.csr_matrix((g[name + '_data'], g[name + '_indices'], g[name + '_indptr']), shape=(n, n))


def test_oracle_assembly_matches_reference_tiny_me0():
    cell, g, p = tiny_cell()
    n = 3 * cell.model.N
    K, M = helpers.oracle_assemble(cell, float(g['omega_top']))
    D, _ = helpers.oracle_assemble(cell, float(g['omega_top']), helpers.oracle_damping_materials(cell))
    for mine, name in ((K, 'K'), (M, 'M'), (D, 'D')):
        ref = _csr(g, name, n)
        assert np.array_equal(mine.indptr, ref.indptr) and np.array_equal(mine.indices, ref.indices)
        assert np.abs(mine.data - ref.data).max() <= 1e-12 * np.abs(ref.data).max()


def test_host_maps_match_reference_tiny_me0():
    cell, g, p = tiny_cell()
    m = cell.model
    fix, zheng, fu, groups = bc_periodic(m, cell.mesh.lines, ndisp=2)
    assert np.array_equal(fix, g['fix']) and np.array_equal(zheng, g['zheng']) and np.array_equal(fu, g['fu'])
    assert np.array_equal(np.array(groups), g['groups'])
    kl = KLayout(m.quads, m.N, 3)
    indptr, indices = kl.csr()
    assert np.array_equal(indptr, g['K_indptr']) and np.array_equal(indices, g['K_indices'])
    bl = BlochLayout(kl, fix, zheng, fu, groups)
    assert np.array_equal(bl.active, g['active'])
    o = np.argsort(g['slaves'])
    assert np.array_equal(bl.slaves, g['slaves'][o]) and np.array_equal(bl.masters, g['masters'][o])
    # reference Km_aa (value-dependent pattern) inside the structural pattern, values via our gathers
    r = np.repeat(np.arange(bl.n_active), np.diff(bl.aa_indptr))
    pos = {(a, b): i for i, (a, b) in enumerate(zip(r.tolist(), bl.aa_indices.tolist()))}
    Kv = g['K_data']
    Kaa = np.array([Kv[bl.aa.idx[bl.aa.ptr[j]:bl.aa.ptr[j + 1]]].sum() for j in range(bl.nnz_aa)])
    ref_slots = np.array([pos[(a, b)] for a, b in zip(g['Kaa_row'].tolist(), g['Kaa_col'].tolist())])
    assert np.abs(Kaa[ref_slots] - g['Kaa_data']).max() <= 1e-12 * np.abs(g['Kaa_data']).max()
    extra = np.setdiff1d(np.arange(bl.nnz_aa), ref_slots)
    assert np.abs(Kaa[extra]).max() <= 1e-12 * np.abs(g['Kaa_data']).max()
    kt = np.array([Kv[bl.kts.idx[bl.kts.ptr[j]:bl.kts.ptr[j + 1]]].sum() for j in range(bl.n_active + 1)])
    x = np.random.default_rng(0).standard_normal(bl.n_active) + 0.5j
    disp = bl.expand(x)
    assert np.isclose(kt[:-1] @ x + kt[-1], g['kts'] @ disp, rtol=1e-12)
