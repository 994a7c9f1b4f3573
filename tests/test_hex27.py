"""Hex27 slab path on the CPU: oracle pins, product host layout and folds, stored curves.

Pins of the oracle (``oracle/fem3d.py``):

* shapes: Kronecker delta at the mesher's node table (ties the node order of the product
  mesher, the kernel table and the oracle's sortSF27 order together), partition of unity;
  quadrature exact to degree 5 per variable with the rounded constants;
* element matrices of a box: tensor products of the 1D quadratic Lagrange stiffness and mass
  matrices (closed form), including the gamma_yz quirk (a C44-only material couples uy-uy by
  d/dz only and leaves uz-uz empty);
* rigid motions: translations and rotations about y and z are in the null space of K_uu,
  the rotation about x is not (the quirk drops duz/dy from gamma_yz);
* K_phiu = KM_SCALE conj(K_uphi)^T under a complex PML stretch; constant potential in the
  null space of K_phiphi; total mass;
* end to end: the reference's stored |Y| curves of the four sp_2p5d slab templates
  (``tests/data/sp_2p5d_*_layer.npz``, copied from the reference's fixtures).
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fem3d
from sawsim_amd import hex27, hexmesh, workloads
from sawsim_amd.layout import BlochLayout, KLayout

DATA = Path(__file__).parent / 'data'


# ------------------------------------------------------------------ shapes and quadrature
def test_h27_shapes_match_the_mesher_node_table():
    N, dN = fem3d.h27_shapes(hexmesh.H27_NODES.astype(float))
    assert np.allclose(N, np.eye(27), atol=1e-15)
    xi = np.random.default_rng(3).uniform(-1, 1, (40, 3))
    N, dN = fem3d.h27_shapes(xi)
    assert np.allclose(N.sum(1), 1.0) and np.allclose(dN.sum(1), 0.0, atol=1e-13)
    f = lambda p: p[..., 0] ** 2 * p[..., 1] ** 2 * p[..., 2] ** 2 - p[..., 0] * p[..., 2] + 2 * p[..., 1]
    assert np.allclose(N @ f(hexmesh.H27_NODES.astype(float)), f(xi))
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        Np, _ = fem3d.h27_shapes(xi + e)
        Nm, _ = fem3d.h27_shapes(xi - e)
        assert np.allclose((Np - Nm) / (2 * h), dN[..., d], atol=1e-8)


def test_h27_gauss_exact_to_degree_5():
    pts, w = fem3d.h27_gauss()
    assert abs(w.sum() - 8.0) < 1e-13
    for a in range(6):
        for b in range(6):
            for c in (0, 2, 4, 5):
                ex = np.prod([2.0 / (k + 1) if k % 2 == 0 else 0.0 for k in (a, b, c)])
                assert abs((w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum() - ex) < 1e-13


def _box(h=(1.3, 0.7, 0.9), origin=(0.2, -0.1, 0.5)):
    pos = (hexmesh.H27_NODES + 1) * 0.5 * np.asarray(h) + np.asarray(origin)
    return pos, np.arange(27)[None, :]


def _lagrange_1d(h):
    """1D quadratic Lagrange stiffness and mass on nodes (0, h/2, h) in order (-1, 0, +1)."""
    K = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)
    M = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30
    return K, M


def _tensor(mats):
    idx = hexmesh.H27_NODES + 1                         # 1D indices per axis
    out = np.ones((27, 27))
    for d in range(3):
        out = out * mats[d][np.ix_(idx[:, d], idx[:, d])]
    return out


def _element(C=None, e=None, eps=None, rho=1.0, pml=None):
    pos, hexas = _box()
    mat = dict(rho=rho, C=np.zeros((6, 6)) if C is None else C, e=np.zeros((3, 6)) if e is None else e,
               eps=np.zeros((3, 3)) if eps is None else eps, indx=np.array([0]))
    if pml is None:
        Ce, ee, pe, re = fem3d.element_tensors(hexas, pos, [mat], [], 1.0)
    else:
        rec = dict(mat, region=3, indx=np.array([0]), **pml)
        Ce, ee, pe, re = fem3d.element_tensors(hexas, pos, [], [rec], 2e9)
    return fem3d.element_matrices(hexas, pos, Ce, ee, pe, re), pos


def test_box_element_matches_tensor_products():
    hx, hy, hz = 1.3, 0.7, 0.9
    (Kx, Mx), (Ky, My), (Kz, Mz) = _lagrange_1d(hx), _lagrange_1d(hy), _lagrange_1d(hz)
    # C11 only: ux-ux couples by d/dx d/dx
    C = np.zeros((6, 6))
    C[0, 0] = 1.0
    (Kuu, _, _, Me), _ = _element(C=C, rho=2.0)
    Kuu = Kuu[0].reshape(27, 3, 27, 3)
    assert np.allclose(Kuu[:, 0, :, 0], _tensor([Kx, My, Mz]), rtol=1e-12, atol=1e-13)
    assert np.allclose(np.delete(np.delete(Kuu.reshape(81, 81), np.arange(0, 81, 3), 0), np.arange(0, 81, 3), 1), 0)
    assert np.allclose(Me[0], 2.0 * _tensor([Mx, My, Mz]), rtol=1e-12, atol=1e-14)
    # C44 only (gamma_yz): only duy/dz enters, so uy-uy gets d/dz d/dz and uz has nothing
    C = np.zeros((6, 6))
    C[3, 3] = 1.0
    (Kuu, _, _, _), _ = _element(C=C)
    Kuu = Kuu[0].reshape(27, 3, 27, 3)
    assert np.allclose(Kuu[:, 1, :, 1], _tensor([Mx, My, Kz]), rtol=1e-12, atol=1e-13)
    assert np.abs(Kuu[:, 2, :, :]).max() < 1e-15 and np.abs(Kuu[:, :, :, 2]).max() < 1e-15
    # eps_xy only: phi-phi couples d/dx with d/dy (mixed tensor product, sign and transpose pinned)
    eps = np.zeros((3, 3))
    eps[0, 1] = 1.0
    (_, _, Kff, _), _ = _element(eps=eps)
    D1 = lambda h: np.array([[-3, 4, -1], [-4, 0, 4], [1, -4, 3]]) / 6.0   # int N_a dN_b/dt over [0, h]
    G = _tensor([D1(hx).T, D1(hy), Mz])
    assert np.allclose(Kff[0], G, rtol=1e-12, atol=1e-13)


def test_rigid_motions_and_the_gamma_yz_quirk():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(6, 6))
    (Kuu, _, _, _), pos = _element(C=A @ A.T)
    K = Kuu[0]
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    vec = lambda ux, uy, uz: np.stack([ux, uy, uz], 1).ravel()
    one, zero = np.ones(27), np.zeros(27)
    scale = np.abs(K).max()
    for u in (vec(one, zero, zero), vec(zero, one, zero), vec(zero, zero, one),
              vec(z, zero, -x), vec(-y, x, zero)):               # translations, rotations about y and z
        assert np.abs(K @ u).max() < 1e-12 * scale
    rot_x = vec(zero, -z, y)                                      # gamma_yz = duy/dz = -1 with the quirk
    assert np.abs(K @ rot_x).max() > 1e-3 * scale


def test_coupling_blocks_under_a_complex_stretch():
    rng = np.random.default_rng(7)
    A = rng.normal(size=(6, 6))
    e = rng.normal(size=(3, 6))
    B = rng.normal(size=(3, 3))
    pml = dict(za=3.0, zp=0.0, xa=0.0, xp=0.0, dmax=1e9, n=3)
    (Kuu, Kuf, Kff, Me), pos = _element(C=A @ A.T, e=e, eps=B @ B.T, rho=3.0, pml=pml)
    assert np.iscomplexobj(Kuf) and np.abs(Kuf.imag).max() > 0
    M, K = fem3d.assemble(np.arange(27)[None, :], pos, [], [dict(region=3, indx=np.array([0]), rho=3.0, C=A @ A.T,
                                                                 e=e, eps=B @ B.T, **pml)], 2e9)
    K = K.toarray()
    ku_f = K[:81, 81:]
    kf_u = K[81:, :81]
    assert np.allclose(kf_u, fem3d.KM_SCALE * ku_f.conj().T, rtol=1e-13, atol=0)
    assert np.abs(K[81:, 81:] @ np.ones(27)).max() < 1e-12 * np.abs(K[81:, 81:]).max()
    # mass: each displacement component carries KM_SCALE rho a1a2a3 * volume
    a3 = 1 + 1e9 * (1 - (pos[26, 2] / 3.0) ** 2) ** 3 / (1j * 2e9)
    vol = 1.3 * 0.7 * 0.9
    Mu = M.toarray()[:81, :81]
    assert np.isclose(Mu[0::3, 0::3].sum(), fem3d.KM_SCALE * 3.0 * a3 * vol, rtol=1e-12)
    assert np.abs(M.toarray()[81:]).max() == 0


# ------------------------------------------------------------------ product host layout and folds
@pytest.fixture(scope='module')
def single():
    cell = workloads.hex_template('single')
    mesh = cell.mesh
    M, K = fem3d.assemble(mesh.hexas, mesh.pos_m, cell.material_list(), cell.pml_list(), 2 * np.pi * 2.7e9)
    m, el = fem3d.boundary_dofs(mesh.sets, mesh.pos_m, 3 * mesh.N)
    return cell, K, M, m, el


def test_template_mesh_counts(single):
    cell = single[0]
    info = cell.mesh.info
    assert info['x_counts'] == [2, 4, 4, 4, 2] and info['z_counts'] == [10, 30, 1]
    assert cell.mesh.ne == 16 * 40 + 8 * 1
    c4 = workloads.config4(points=3)
    assert 7500 <= c4.mesh.ne <= 8500


def test_boundary_sets_and_fold_match_the_oracle(single):
    cell, K, M, m, el = single
    mesh = cell.mesh
    bnd = hex27.boundary_idx_vect_3ds(hex27.find_saw_boundary_3ds(mesh), mesh.pos_m, 3 * mesh.N)
    for k in m:
        assert np.array_equal(bnd.machine[k], m[k]), k
    for k in el:
        assert np.array_equal(bnd.electric[k], el[k]), k
    gdof = 4 * mesh.N
    S = fem3d.Solver3D(K, M, m, el)
    rep = hex27.fold_map(bnd, gdof)
    assert np.array_equal(rep, hex27.rep_from_operators(S.t_fb, S.t_lr))
    assert np.array_equal(np.setdiff1d(np.arange(gdof), hex27.prescribed_dofs(bnd)), S.active)
    # paired coordinates: L/R differ by the period in x, front/back by the aperture in y
    L, R = bnd.machine['period_L'] // 3, bnd.machine['period_R'] // 3
    assert np.allclose(mesh.pos_m[R, 0] - mesh.pos_m[L, 0], mesh.info['period_m'], atol=1e-15)
    assert np.allclose(mesh.pos_m[R][:, 1:], mesh.pos_m[L][:, 1:], atol=1e-15)


def test_gathered_active_system_matches_the_two_stage_fold(single):
    """The product's slot layout + Bloch gathers, run on the CPU from the oracle's K and M,
    reproduce T_lr^T (T_fb^T A T_fb) T_lr restricted to the active set, the RHS column and the
    charge rows."""
    cell, K, M, m, el = single
    mesh = cell.mesh
    kl = KLayout(mesh.hexas, mesh.N, 4, numbering='interleaved')
    rows, cols = kl.slot_rows_cols()
    Kd, Md = K.tocsr(), M.tocsr()
    kv = np.asarray(Kd[rows, cols]).ravel()
    mv = np.asarray(Md[rows, cols]).ravel()
    assert np.abs(Kd).sum() == pytest.approx(np.abs(kv).sum(), rel=1e-13)      # every entry has a slot
    bnd = hex27.boundary_idx_vect_3ds(mesh.sets, mesh.pos_m, 3 * mesh.N)
    bl = BlochLayout(kl, bnd.machine['fix'], bnd.electric['zheng'], bnd.electric['fu'],
                     rep=hex27.fold_map(bnd, kl.gdof))
    S = fem3d.Solver3D(K, M, m, el)
    assert np.array_equal(bl.active, S.active)

    def gather(gl, v):
        out = np.zeros(gl.ptr.size - 1, dtype=complex)
        np.add.at(out, np.repeat(np.arange(gl.ptr.size - 1), np.diff(gl.ptr)), v[gl.idx])
        return out
    import scipy.sparse as sp
    Kaa = sp.csr_matrix((gather(bl.aa, kv), bl.aa_indices, bl.aa_indptr), shape=(bl.n_active,) * 2)
    ref = S.k_aa.tocsr()
    assert abs(Kaa - ref).max() <= 1e-13 * abs(ref).max()
    # pattern: every stored reference entry is a slot of the product pattern
    rr = ref.tocoo()
    pat = sp.csr_matrix((np.ones(bl.aa_indices.size), bl.aa_indices, bl.aa_indptr), shape=(bl.n_active,) * 2)
    assert np.all(np.asarray(pat[rr.row, rr.col]).ravel() == 1)
    rK = gather(bl.rhs, kv)
    assert np.allclose(rK, np.asarray(S.k_az.sum(axis=1)).ravel(), rtol=0, atol=1e-13 * abs(rK).max())
    kts = gather(bl.kts, kv)
    x = np.random.default_rng(2).normal(size=bl.n_active) + 1j
    q_ref = (S.kz_rows @ S.expand(x, 1.0)).sum()
    assert abs(kts[:-1] @ x + kts[-1] - q_ref) <= 1e-12 * abs(q_ref)


# ------------------------------------------------------------------ stored reference curves
def _refined_solve(S, om):
    A = (S.k_aa - om ** 2 * S.m_aa).tocsc()
    rhs = -np.asarray((S.k_az - om ** 2 * S.m_az).sum(axis=1)).ravel()
    lu = spla.splu(A)
    x = lu.solve(rhs)
    for _ in range(2):
        x = x + lu.solve(rhs - A @ x)
    return (S.kz_rows @ S.expand(x)).sum()


@pytest.mark.parametrize('name', ['single', 'quad'])
def test_oracle_reproduces_stored_slab_curves(name):
    """|i w Q| (total slice admittance) of the oracle on the rebuilt slab vs the reference's
    stored solver curve.  The reference's own regression tolerance is 2e-3; the stored values
    come from a single direct solve of a system whose first SuperLU solve moves Q by ~2e-6 under
    refinement, so agreement is bounded by that solver noise, not by the model."""
    meta = json.loads((DATA / f'sp_2p5d_{name}_layer.json').read_text())
    d = np.load(DATA / f'sp_2p5d_{name}_layer.npz')
    f_all = d['frequency_hz']
    cell = workloads.hex_template(name)
    mesh = cell.mesh
    M, K = fem3d.assemble(mesh.hexas, mesh.pos_m, cell.material_list(), cell.pml_list(), 2 * np.pi * f_all[-1])
    m, el = fem3d.boundary_dofs(mesh.sets, mesh.pos_m, 3 * mesh.N)
    S = fem3d.Solver3D(K, M, m, el)
    for i in meta['test_indices']:
        om = 2 * np.pi * f_all[i]
        y = abs(1j * om * _refined_solve(S, om))
        own = d['sawsim_magnitude'][i]
        assert abs(y - own) / own <= 5e-5
        ref = d['reference_magnitude'][i]
        assert abs(y - ref) / ref <= meta['reference_rtol']


# ------------------------------------------------------------------ reference golden vectors
def hex27_golden():
    """tests/golden/golden_hex27_tiny.npz (written by tests/golden/make_golden.py from the
    reference's saw3d modules) unpacked into meshes, records and expected outputs."""
    g = np.load(Path(__file__).parent / 'golden' / 'golden_hex27_tiny.npz')
    import scipy.sparse as sp

    def recs(prefix, n):
        out = []
        for i in range(n):
            keys = [k for k in g.files if k.startswith(f'{prefix}{i}_')]
            out.append({k[len(f'{prefix}{i}_'):]: (g[k].item() if g[k].ndim == 0 else g[k]) for k in keys})
        return out
    nsub = g['sub_coord'].shape[0]
    csr = lambda p: sp.csr_matrix((g[p + '_data'], g[p + '_indices'], g[p + '_indptr']), shape=(4 * nsub, 4 * nsub))
    sets = {k[len('slab_sets_'):]: g[k] for k in g.files if k.startswith('slab_sets_')}
    bnd = {part: {k[len(f'bnd_{part}_'):]: g[k] for k in g.files if k.startswith(f'bnd_{part}_')}
           for part in ('machine', 'electric')}
    return g, recs('sub_mat', 2), recs('sub_pml', 3), csr('sub_K'), csr('sub_M'), sets, bnd


def test_oracle_assembly_matches_reference_golden_with_x_z_and_xz_stretch():
    """K and M of the reference's form_stiffness_mass_3d on six elements with PML records of
    region 3 (z), 5 (x) and 2 (x and z) stretch, each with a rotated 3m record: the oracle's
    entries within 1e-12 of the largest, on the same pattern."""
    g, mats, pmls, Kr, Mr, _, _ = hex27_golden()
    M, K = fem3d.assemble(g['sub_elem'], g['sub_coord'], mats, pmls, float(g['sub_omega']))
    for ours, ref in ((K, Kr), (M, Mr)):
        d = (ours.tocsr() - ref).tocsr()
        assert abs(d).max() <= 1e-12 * abs(ref).max()
        assert (abs(ref) > 0).sum() == (abs(ours) > 0).sum()


def test_boundary_sets_and_fold_match_reference_golden():
    """The reference's boundary_idx_vect_3ds sets and its two fold operators on the 84-element
    slab: the product's host sets and the oracle's equal them element for element, and the
    product's representative map equals the composite fold T_fb T_lr."""
    g, _, _, _, _, sets, bnd = hex27_golden()
    pos = g['slab_pos']
    N = pos.shape[0]
    b = hex27.boundary_idx_vect_3ds(sets, pos, 3 * N)
    m, el = fem3d.boundary_dofs(sets, pos, 3 * N)
    for part, ours, orc in (('machine', b.machine, m), ('electric', b.electric, el)):
        assert set(ours) == set(bnd[part]) == set(orc)
        for k, ref in bnd[part].items():
            assert np.array_equal(np.asarray(ours[k]), ref) and np.array_equal(np.asarray(orc[k]), ref), (part, k)
    gd = 4 * N
    rep = hex27.fold_map(b, gd)
    t_fb, t_lr = g['t_fb'], g['t_lr']
    assert np.array_equal(rep, t_lr[t_fb])             # one 1 per row: (T_fb T_lr)[d] = lr(fb(d))
    assert np.array_equal(np.setdiff1d(np.arange(gd), hex27.prescribed_dofs(b)), g['slab_active'])


def test_oracle_slab_charge_matches_reference_golden():
    """Q of the reference's FrequencySolver3D (SuperLU + 3 refinement steps on its own system)
    at three frequencies: the oracle, on its own assembly and folds, within 1e-9."""
    g, _, _, _, _, sets, _ = hex27_golden()
    mats = workloads.hex_record_materials(tuple(g['slab_euler']), 'al', [])
    mesh = hexmesh.template_slab('single', 1.085, 0.217, 0.17, 1.2, 0.5, 0.5, [])
    assert np.array_equal(mesh.pos_m, g['slab_pos']) and np.array_equal(mesh.hexas, g['slab_hexas'])
    cell = workloads.HexCell(mesh=mesh, materials=mats, band_hz=(1.8e9, 2.7e9), points=3)
    M, K = fem3d.assemble(mesh.hexas, mesh.pos_m, cell.material_list(), cell.pml_list(), float(g['slab_omega_top']))
    assert K.nnz >= int(g['slab_K_nnz'])
    m, el = fem3d.boundary_dofs(mesh.sets, mesh.pos_m, 3 * mesh.N)
    eng = fem3d.Solver3D(K, M, m, el)
    for w, qr in zip(2 * np.pi * g['slab_frequency_hz'], g['slab_q_refined']):
        a = (eng.k_aa - w * w * eng.m_aa).tocsc()
        rhs = -np.asarray((eng.k_az - w * w * eng.m_az).sum(axis=1)).ravel()
        lu = spla.splu(a)
        x = lu.solve(rhs)
        for _ in range(3):
            x = x + lu.solve(rhs - a @ x)
        q = (eng.kz_rows @ eng.expand(x)).sum()
        assert abs(q - qr) <= 1e-9 * abs(qr), (q, qr)
