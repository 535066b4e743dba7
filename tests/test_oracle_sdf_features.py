"""The oracle's edge and triangle contacts against analytic spheres and grid
SDFs (Physics/ClothCollision.cpp:472-579 with ClothCollisionSDF.h:26-55,
148-205), pinned by closed forms."""
import numpy as np
import pytest

CF_VERTICES, CF_TRIANGLES, CF_EDGES = 1, 2, 16


def cloth(oracle_mod, pos, tris, thickness, tolerance, flags):
    pos = np.asarray(pos, dtype=np.float64)
    c = oracle_mod.Cloth()
    c.set(thickness=thickness, tolerance=tolerance, collFlags=flags, friction=0.0, fem=1)
    c.init_arrays(pos, pos[:, [0, 2]], np.ones(len(pos)), np.asarray(tris, dtype=np.int32))
    return c


def test_sphere_segment_query(oracle_mod):
    """Edge above a sphere: closest segment point, radial normal, surface
    point, weights (1 - t, t)."""
    r = 0.5
    pos = np.array([[-0.3, 0.51, 0.1], [0.5, 0.51, 0.1], [0.0, 2.0, 0.0]])
    c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=CF_EDGES)
    c.add_sphere((0.0, 0.0, 0.0), r)
    c.detect()
    ids, data = c.edge_contacts()
    assert len(ids) == 1 and ids[0, 2] == 0 and ids[0, 3] == -1
    t = 0.3 / 0.8
    q = np.array([0.0, 0.51, 0.1])
    n = q / np.linalg.norm(q)
    np.testing.assert_allclose(data[0, 3:6], n, atol=1e-6)
    np.testing.assert_allclose(data[0, :3], r * n, atol=1e-6)
    np.testing.assert_allclose(data[0, 6:8], [1 - t, t], atol=1e-6)


@pytest.mark.parametrize("flags,expect", [(CF_EDGES, 1), (CF_EDGES | CF_VERTICES, 0)])
def test_sphere_edge_border_filter(oracle_mod, flags, expect):
    """An edge whose closest point is its end point (w = (1, 0)) is kept only
    when vertex contacts are off (handleBorder, ClothCollision.cpp:474)."""
    pos = np.array([[0.0, 0.505, 0.0], [0.5, 0.9, 0.0], [0.0, 2.0, 0.5]])
    c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=flags)
    c.add_sphere((0.0, 0.0, 0.0), 0.5)
    c.detect()
    ids, data = c.edge_contacts()
    got = [i for i, d in zip(ids, data) if i[2] == 0]
    assert len(got) == expect


def test_sphere_triangle_query(oracle_mod):
    """Centre below a triangle's interior: barycentric coordinates of the
    projection, normal towards the triangle, distance = height - radius."""
    pos = np.array([[-0.2, 0.505, -0.1], [0.3, 0.505, -0.1], [-0.2, 0.505, 0.4]])
    c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=CF_TRIANGLES)
    c.add_sphere((0.0, 0.0, 0.0), 0.5)
    c.detect()
    ids, data = c.tri_contacts()
    assert len(ids) == 1 and ids[0, 1] == -1
    # projection (0, 0.505, 0) in the (x, z) triangle a=(-0.2,-0.1) b=(0.3,-0.1) c=(-0.2,0.4)
    exp = np.array([0.4, 0.4, 0.2])
    np.testing.assert_allclose(data[0, 6:9], exp, atol=1e-6)
    np.testing.assert_allclose(data[0, 3:6], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(data[0, :3], [0, 0.5, 0], atol=1e-6)


def linear_sdf(oracle_mod, c, n=8, half=1.0):
    """f(p) = p.y sampled on an n^3-cell grid over [-half, half]^3."""
    bmin, bmax = (-half,) * 3, (half,) * 3
    step = np.float32(2 * half) * (np.float32(1) / np.float32(n))
    idx = np.arange(n + 1, dtype=np.float32)
    Z, Y, X = np.meshgrid(idx, idx, idx, indexing="ij")
    vals = (Y * step + np.float32(-half)).astype(np.float32).ravel()
    c.add_grid_sdf(vals, (n, n, n), bmin, bmax)


def frank_wolfe_linear(a, b, c):
    """Closed form of the Frank-Wolfe iterate for f = y: the target vertex is
    the lowest one, cp_k - s = (cp_0 - s) * 2 / ((k + 1)(k + 2))."""
    verts = [a, b, c]
    s = verts[int(np.argmin([v[1] for v in verts]))]
    cp0 = 0.33 * (a + b + c)
    dist_old = 0.0
    cp = cp0
    for i in range(1, 100):
        dist = cp[1]
        if dist_old > 0 and abs(dist_old - dist) / dist_old < 1e-4:
            break
        dist_old = dist
        cp = s + (cp0 - s) * 2.0 / ((i + 1) * (i + 2))
    return cp


def test_grid_sdf_triangle_frank_wolfe(oracle_mod):
    """GridSDF::QueryTriangle on a linear field: the iterate converges to the
    lowest vertex along the closed-form Frank-Wolfe path; the contact point
    is cp - f(cp) grad, the coordinates those of cp."""
    a = np.array([-0.2, 0.004, -0.1])
    b = np.array([0.3, 0.03, -0.1])
    cc = np.array([-0.2, 0.05, 0.4])
    pos = np.array([a, b, cc])
    c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=CF_TRIANGLES)
    linear_sdf(oracle_mod, c)
    c.detect()
    ids, data = c.tri_contacts()
    cp = frank_wolfe_linear(a, b, cc)
    assert cp[1] < 0.03  # within thickness + tolerance
    assert len(ids) == 1
    np.testing.assert_allclose(data[0, :3], [cp[0], 0.0, cp[2]], atol=1e-5)
    np.testing.assert_allclose(data[0, 3:6], [0, 1, 0], atol=1e-6)
    # barycentric of cp: mostly the lowest vertex a
    w = data[0, 6:9]
    np.testing.assert_allclose(w[0] * a + w[1] * b + w[2] * cc, cp, atol=1e-5)


def test_grid_sdf_triangle_border_filter(oracle_mod):
    """With edge contacts on, a triangle contact whose coordinates touch the
    border (< 0.001) is dropped; with them off it is kept."""
    a = np.array([-0.2, 0.004, -0.1])
    b = np.array([0.3, 0.3, -0.1])
    cc = np.array([-0.2, 0.3, 0.4])
    pos = np.array([a, b, cc])
    kept = []
    for flags in (CF_TRIANGLES, CF_TRIANGLES | CF_EDGES | CF_VERTICES):
        c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=flags)
        linear_sdf(oracle_mod, c)
        c.detect()
        ids, data = c.tri_contacts()
        kept.append(len(ids))
        if len(ids):
            assert data[0, 7] < 0.001 or data[0, 8] < 0.001
    assert kept == [1, 0]


def test_grid_sdf_edges_need_vertex_contacts(oracle_mod):
    pos = np.array([[-0.2, 0.004, -0.1], [0.3, 0.3, -0.1], [-0.2, 0.3, 0.4]])
    c = cloth(oracle_mod, pos, [[0, 1, 2]], thickness=0.01, tolerance=0.02, flags=CF_EDGES)
    linear_sdf(oracle_mod, c)
    with pytest.raises(RuntimeError):
        c.step(1.0 / 60)
    c.set(collFlags=CF_EDGES | CF_VERTICES)
    c.detect()
    assert len(c.edge_contacts()[0]) == 0
