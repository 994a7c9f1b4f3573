"""Parameter -> tensor maps and their Jacobians (CPU)."""
import numpy as np
import pytest

from sawsim_amd import materials as mt, params as prm, workloads


@pytest.mark.parametrize('cfg', ['config1', 'config3'])
def test_lifted_tensors_match_kernel_inputs(cfg):
    cell = getattr(workloads, cfg)(mesh_um=0.3, points=3)
    C9, E9, EP, _ = mt.kernel_tensors([mt.lift_projected(cell.materials[0])])
    C, E, P = prm.lifted_tensors(cell.params, 'xy')
    assert np.allclose(C, C9[0], rtol=1e-13, atol=1e-13 * np.abs(C9).max())
    assert np.allclose(E, E9[0], rtol=1e-13, atol=1e-13)
    assert np.allclose(P, EP[0], rtol=1e-13, atol=1e-25)


def test_complex_step_jacobian_matches_central_differences():
    p = workloads.config3(mesh_um=0.3, points=3).params.copy()
    p[12] = 3.0   # nonzero alpha
    dC, dE, dP = prm.jacobian(p)
    for k in range(17):
        h = 1e-6 * abs(p[k]) if p[k] else 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        Cp, Ep, Pp = prm.lifted_tensors(pp)
        Cm, Em, Pm = prm.lifted_tensors(pm)
        if k == 15:        # beta_dk does not enter the tensors
            continue
        for got, a, b in ((dC[k], Cp, Cm), (dE[k], Ep, Em), (dP[k], Pp, Pm)):
            fd = (a - b) / (2 * h)
            # central differences carry rounding noise ~1e-16 |tensor| / h
            assert np.allclose(got, fd, rtol=1e-5, atol=1e-8 * np.abs(a).max() / h)


def test_full_tensors_match_the_slab_piezo_and_their_jacobian():
    """The 18-parameter slab Jacobian: tensors equal the assembled piezo record (4th-rank
    rotation route), derivatives equal central differences, density enters only rho."""
    from sawsim_amd import workloads
    cell = workloads.config4(mesh_um=0.3, points=3, beta_dk=1e-13, eta_eps=1e-3)
    p = cell.params
    assert p.size == 18 and prm.NAMES3[17] == 'rho'
    C, e, eps = prm.full_tensors(p)
    pz = cell.materials[0]
    assert np.allclose(C, pz.C, rtol=1e-12, atol=1e-12 * np.abs(pz.C).max())
    assert np.allclose(e, pz.e, rtol=1e-12, atol=1e-12 * np.abs(pz.e).max())
    assert np.allclose(eps, pz.eps, rtol=1e-12, atol=1e-12 * np.abs(pz.eps).max())
    dC, dE, dP, drho = prm.jacobian_full(p)
    for k in range(17):
        h = 1e-6 * abs(p[k]) if p[k] else 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        Cp, ep, sp_ = prm.full_tensors(pp)
        Cm, em, sm = prm.full_tensors(pm)
        for got, a, b in ((dC[k], Cp, Cm), (dE[k], ep, em), (dP[k], sp_, sm)):
            fd = (a - b) / (2 * h)
            assert np.abs(got - fd).max() <= 1e-6 * max(np.abs(fd).max(), 1e-30) + 1e-12 * np.abs(a).max() / max(abs(p[k]), 1)
    assert np.count_nonzero(drho) == 1 and drho[17] == 1
