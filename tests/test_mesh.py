"""Structured Q9 mesher: node counts of the reference templates and sp_stack KATs."""
import numpy as np
import pytest

from sawsim_amd.mesh import double_layer_cell, quad_layer_cell, stack_cell


def test_template_element_and_node_counts():
    # sp_meshes/double.py:13 (340 elements), sp_meshes/quad.py docstring (372 elements, 1601 nodes)
    assert double_layer_cell().ne == 340
    q = quad_layer_cell()
    assert (q.ne, q.N) == (372, 1601)


@pytest.mark.parametrize('piezo,backings,nodes', [
    (8.0, [], 1917),                     # sp_single_layer stack: 7668 DOF = 4 x 1917
    (0.6, [8.33], 2115),                 # sp_double_layer stack: 8460 DOF
    (0.6, [0.5, 6.9], 1983),             # sp_triple_layer stack: 7932 DOF
    (0.6, [0.5, 1.0, 7.83], 2247),       # sp_quad_layer stack: 8988 DOF
])
def test_sp_stack_dof_kats(piezo, backings, nodes):
    # docs/models.en.md:60-66 at the converge mesh (mesh_um = 0.25)
    assert stack_cell(1.085, 0.5, 0.17, piezo, backings, [], 0.25).N == nodes


def test_q9_connectivity_is_conforming_and_ccw():
    m = stack_cell(1.085, 0.5, 0.0868, 0.6, [4.825], [0.3255], 0.1, pml_pitches=1.0)
    P, Q = m.pos_m, m.quads
    # counter-clockwise corners (positive signed area) and midpoints at edge centres
    a = P[Q[:, 0]]; b = P[Q[:, 1]]; c = P[Q[:, 2]]; d = P[Q[:, 3]]
    area = 0.5 * ((b[:, 0] - a[:, 0]) * (d[:, 1] - a[:, 1]) - (d[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))
    assert (area > 0).all()
    assert np.allclose(P[Q[:, 4]], 0.5 * (a + b)) and np.allclose(P[Q[:, 8]], 0.25 * (a + b + c + d))
    # every node is used, no duplicate coordinates
    assert np.unique(Q).size == m.N
    assert np.unique(np.round(P * 1e12).astype(np.int64), axis=0).shape[0] == m.N


def test_periodic_boundaries_pair_and_electrodes():
    m = stack_cell(1.085, 0.5, 0.0868, 0.6, [4.825], [], 0.1, pml_pitches=1.0)
    L, R = m.nodes_with_tag(19), m.nodes_with_tag(20)
    assert L.size == R.size
    assert np.allclose(np.sort(m.pos_m[L, 1]), np.sort(m.pos_m[R, 1]), atol=1e-14)
    assert np.allclose(m.pos_m[R, 0] - 2.17e-6, 0.0, atol=1e-15)
    z = m.nodes_with_tag(17)
    assert np.intersect1d(z, np.concatenate([L, R])).size == 0
    # electrode + boundary: bottom, top, two sides of strips 1-2
    x = m.pos_m[z, 0]
    assert np.isclose(x.min(), 1.085e-6 * 0.25) and np.isclose(x.max(), 1.085e-6 * 0.75)
