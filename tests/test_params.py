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
