"""Case geometry: lattice counts and wall normals of the dam-break cases."""
import numpy as np

from sphinxsys_amd import cases, geometry


def test_dambreak_2d_particle_counts():
    g = cases.dambreak_2d(order="lattice")
    assert len(g.fluid_pos) == 80 * 40
    assert len(g.wall_pos) == 3504
    assert np.allclose(g.fluid_pos.min(axis=0), [0.0125, 0.0125])
    assert np.allclose(g.fluid_pos.max(axis=0), [1.9875, 0.9875])


def test_wall_normals_point_from_wall_into_the_tank():
    g = cases.dambreak_2d()
    # innermost layer of the bottom wall: y = -dp/2, normal +y
    bottom = np.isclose(g.wall_pos[:, 1], -0.0125) & (g.wall_pos[:, 0] > 0.5) & (g.wall_pos[:, 0] < 4.5)
    assert bottom.sum() > 100
    assert np.allclose(g.wall_normal[bottom], [0.0, 1.0], atol=1e-12)
    # lattice columns are at -0.1875 + 0.025 k: the innermost right-wall column is x = 5.3875
    right = np.isclose(g.wall_pos[:, 0], 5.3875, atol=1e-9) & (g.wall_pos[:, 1] > 0.5) & (g.wall_pos[:, 1] < 4.5)
    assert right.sum() > 100
    assert np.allclose(g.wall_normal[right], [-1.0, 0.0], atol=1e-12)
    assert np.allclose(np.linalg.norm(g.wall_normal, axis=1), 1.0)


def test_box_closest_point_inside_projects_to_nearest_face():
    b = geometry.Box((0.0, 0.0, 0.0), (1.0, 2.0, 3.0))
    p = np.array([[0.9, 0.0, 0.0], [0.0, -1.95, 0.5], [5.0, 0.0, 0.0]])
    c = b.closest_point(p.copy())
    assert np.allclose(c, [[1.0, 0.0, 0.0], [0.0, -2.0, 0.5], [1.0, 0.0, 0.0]])


def test_dambreak_3d_small_counts():
    g = cases.dambreak_3d(dp=0.05)
    n = len(g.fluid_pos)
    assert n == 24 * 20 * 12
    assert g.wall_normal.shape == g.wall_pos.shape


def test_morton_order_is_a_permutation_sorted_by_cell_key():
    g = cases.dambreak_2d(order="lattice")
    m = cases.dambreak_2d(order="morton")
    assert np.array_equal(np.sort(g.fluid_pos, axis=0), np.sort(m.fluid_pos, axis=0))
    perm = geometry.morton_order(g.fluid_pos, g.bounds_lower, 2.6 * g.dp)
    assert np.array_equal(m.fluid_pos, g.fluid_pos[perm])
