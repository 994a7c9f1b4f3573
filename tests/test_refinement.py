"""Stopping rule of the iterative refinement (host logic, CPU)."""
import numpy as np

from sawsim_amd.sweep import ir_error


def test_ir_error_is_exact_for_a_geometric_sequence():
    # Q_k = Q* + c r^k: the estimate after step k equals the true remaining error |Q_k - Q*|
    qs = 2.0 + 0.5j
    for r in (1e-3, 0.05, 0.3):
        c = 1e-4 * (1 + 1j)
        q = [qs + c * r ** k for k in range(5)]
        prev = np.inf
        for k in range(1, 5):
            dq = abs(q[k] - q[k - 1]) / abs(q[k])
            est = float(ir_error(dq, prev))
            true = abs(q[k] - qs) / abs(q[k])
            if k == 1:
                assert est == dq            # no rate yet: the correction itself
            else:
                assert np.isclose(est, true, rtol=1e-6)
            prev = dq


def test_ir_error_falls_back_to_the_correction_when_contraction_is_slow():
    assert float(ir_error(1e-8, 1.5e-8)) == 1e-8     # rate 0.67: no extrapolation
    assert float(ir_error(1e-8, 1e-8)) == 1e-8       # stalled
    est = ir_error(np.array([1e-7, 1e-7]), np.array([1e-4, np.inf]))
    assert np.isclose(est[0], 1e-10 / (1 - 1e-3)) and est[1] == 1e-7
