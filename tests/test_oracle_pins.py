"""Pins of oracle functions on the drape path by closed forms, each built so
that a plausible mistake changes the answer (named in each test):

* GridSDF gradients (Geometry/SDF.cpp:210-256): forward differences with
  the last layer copied, on a quadratic field where forward and central
  differences differ;
* SDF / sphere vertex-contact detection (Physics/ClothCollision.cpp:438-470):
  the query from the previous position, the radius thickness +
  max(tolerance, 1.1 |prev - pos|), and 0 <= distance;
* the reference's early exit (Physics/ClothPBD.cpp:324-369): the
  vertex-contact threshold 0.01, the bend threshold 0.005 and the bend
  frequency, by iteration counts of constructed cases.
"""
import math

import numpy as np
import pytest

CF_VERTICES, CF_WALLS = 1, 8


# ------------------------------------------------------------ SDF gradients

def test_sdf_gradient_forward_differences_on_a_quadratic(oracle_mod):
    """f = x^2 + 2 y^2 + 3 z^2 + x y - y z on a grid of step 1/2 (all node
    values, differences and divisions exact in FP32). Forward differences
    give df/dx = 2x + y + h, df/dy = 4y + x - z + 2h, df/dz = 6z - y + 3h at
    nodes below the last layer of each axis; central differences would drop
    the h terms. The last layer of x copies layer nx-1, then (per x) the last
    z layer copies nz-1 and the last y layer copies ny-1 (corners included):
    a scheme that extrapolated, or copied from the wrong neighbour, fails."""
    nx, ny, nz = 4, 5, 6
    h = 0.5
    c = oracle_mod.Cloth(real64=False)
    i = np.arange(nx + 1) * h
    j = np.arange(ny + 1) * h
    k = np.arange(nz + 1) * h
    Z, Y, X = np.meshgrid(k, j, i, indexing="ij")
    F = X ** 2 + 2 * Y ** 2 + 3 * Z ** 2 + X * Y - Y * Z
    c.add_grid_sdf(F.astype(np.float32).ravel(), (nx, ny, nz), (0, 0, 0), (nx * h, ny * h, nz * h))
    G = c.sdf_grads(0).reshape(nz + 1, ny + 1, nx + 1, 3)
    gx = 2 * X + Y + h
    gy = 4 * Y + X - Z + 2 * h
    gz = 6 * Z - Y + 3 * h
    want = np.stack([gx, gy, gz], axis=-1)
    # the copies: x last layer from nx-1, z last layer from nz-1, y from ny-1
    want[:, :, nx] = want[:, :, nx - 1]
    want[nz, :, :nx] = want[nz - 1, :, :nx]
    want[:, ny, :nx] = want[:, ny - 1, :nx]
    # x = nx copies the (already copied) layer nx-1 in full
    want[:, :, nx] = want[:, :, nx - 1]
    assert np.array_equal(G, want)
    # central differences at an interior node differ by the h terms
    assert G[2, 2, 2, 0] - (2 * X[2, 2, 2] + Y[2, 2, 2]) == h


# ------------------------------------------------ SDF vertex-contact rule

THICK, TOL = 0.01, 0.02


def _plane_sdf_cloth(oracle_mod, prev_y, pos_y):
    """Particles far apart in x over the grid SDF of the plane f = y (node
    values exact), one triangle per three particles; prev and pos set."""
    n = len(prev_y)
    c = oracle_mod.Cloth(real64=False)
    c.set(unit=1, fem=1, thickness=THICK, tolerance=TOL, collFlags=CF_VERTICES)
    cells = 8
    z, y, x = np.meshgrid(*(np.linspace(-1, 1, cells + 1),) * 3, indexing="ij")
    c.add_grid_sdf(y.astype(np.float32).ravel(), (cells,) * 3, (-1, -1, -1), (1, 1, 1))
    xs = -0.75 + 0.125 * np.arange(n)
    prev = np.stack([xs, np.asarray(prev_y, float), np.zeros(n)], axis=1)
    pos = np.stack([xs, np.asarray(pos_y, float), np.zeros(n)], axis=1)
    m = n - n % 3
    tris = np.arange(m).reshape(-1, 3)
    c.init_arrays(prev, prev[:, [0, 2]], np.ones(n), tris)
    c.put("prev", prev)
    c.put("pos", pos)
    return c


def test_sdf_vertex_detection_rule(oracle_mod):
    """Grid-SDF vertex contacts, ClothCollision.cpp:438-470. With distance
    d measured at prev and radTol = thickness + max(tolerance, 1.1 |prev -
    pos|) (thickness 0.01, tolerance 0.02), a contact exists iff 0 <= d <
    radTol. Case by case (prev.y, pos.y -> contact?), and the mistake each
    one catches:
      (0.025, 0.025) yes: radTol 0.03 from the tolerance (a radius without it
                          would be 0.01);
      (0.04, 0.04)   no:  beyond radTol 0.03;
      (0.115, 0.015) yes: |prev - pos| = 0.1, radTol 0.12 with the 1.1 factor
                          (0.11 without it);
      (0.07, 0.02)   no:  |prev - pos| = 0.05, radTol 0.065 < 0.07; queried
                          at pos (0.02) it would be a contact;
      (-0.005, 0.3)  no:  d < 0 (dropping the 0 <= d bound would add it).
    Each contact's point is prev - d n and its normal the field's (0, 1, 0)."""
    prev_y = [0.025, 0.04, 0.115, 0.07, -0.005, 0.5]
    pos_y = [0.025, 0.04, 0.015, 0.02, 0.3, 0.5]
    want = {0: True, 1: False, 2: True, 3: False, 4: False, 5: False}
    c = _plane_sdf_cloth(oracle_mod, prev_y, pos_y)
    c.detect()
    ids, data = c.contacts()
    got = {int(i): True for i in ids[:, 0]}
    assert {k: got.get(k, False) for k in want} == want
    prev = c.get("prev")
    for (v, coll), d in zip(ids, data):
        np.testing.assert_allclose(d[3:6], [0, 1, 0], atol=1e-6)
        np.testing.assert_allclose(d[0:3], [prev[v, 0], 0.0, prev[v, 2]], atol=2e-7)


def test_sphere_vertex_detection_rule(oracle_mod):
    """The same rule through the analytic sphere (ClothCollisionSDF.h:8-24):
    d = |prev - c| - r, normal (prev - c) / |prev - c|, point c + r n."""
    c = oracle_mod.Cloth(real64=False)
    c.set(unit=1, fem=1, thickness=THICK, tolerance=TOL, collFlags=CF_VERTICES)
    c.add_sphere((0.0, 0.0, 0.0), 0.5)
    # radial distances above the surface at prev, and the travel
    cases = [(0.025, 0.0, True), (0.04, 0.0, False), (0.115, 0.1, True), (0.14, 0.115, False),
             (-0.005, 0.2, False)]
    dirs = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1], [-1, 0, 0], [0, 0, -1], [0, -1, 0]], float)
    prev, pos = [], []
    for (d, travel, _), u in zip(cases, dirs):
        prev.append(u * (0.5 + d))
        pos.append(u * (0.5 + d - travel) if d > 0 else u * (0.5 + d + travel))
    prev.append(dirs[5] * 2.0)
    pos.append(dirs[5] * 2.0)
    prev, pos = np.array(prev), np.array(pos)
    c.init_arrays(prev, prev[:, [0, 2]], np.ones(6), np.arange(6).reshape(2, 3))
    c.put("prev", prev)
    c.put("pos", pos)
    c.detect()
    ids, data = c.contacts()
    got = sorted(int(i) for i in ids[:, 0])
    assert got == [k for k, cs in enumerate(cases) if cs[2]]
    for (v, coll), d in zip(ids, data):
        u = dirs[v]
        np.testing.assert_allclose(d[3:6], u, atol=1e-6)
        np.testing.assert_allclose(d[0:3], 0.5 * u, atol=1e-6)


# ------------------------------------------------------------ early exit

G = 9.8


def _exit_cloth(oracle_mod, **kw):
    c = oracle_mod.Cloth(real64=False)
    c.set(unit=1, fem=1, dihedral=1, iterations=50, fixedIterations=0, friction=0.0, numSteps=1, **kw)
    return c


def _cancel_gravity(c, h, n):
    """Initial velocity +g h: the prediction leaves every position in place."""
    c.put("vel", np.tile([0.0, G * h, 0.0], (n, 1)))


def test_flat_fem_cloth_exits_after_one_iteration(oracle_mod):
    """FEM mode never sets the link error and a flat cloth has no bend
    error, so the reference's loop stops after its first pass
    (ClothPBD.cpp:337-358). An exit test that also waited for the FEM
    error, or counted from 0, would give another count."""
    h = 1.0 / 240
    c = _exit_cloth(oracle_mod, bend=1.0)
    c.init_grid(6, 6, 0.1, (0, 0, 0), True, False)
    _cancel_gravity(c, h, 36)
    c.substep(h)
    assert c.stats()["iterations"] == 1


@pytest.mark.parametrize("below,iterations", [(-0.002, 1), (0.002, 2)])
def test_contact_exit_threshold(oracle_mod, below, iterations):
    """A flat cloth on the ground plane y = -1 with thickness 0.01:
    y = -1 + 0.002 leaves a contact error of
    0.008 < 0.01, so the first pass ends the loop; y = -1 - 0.002 leaves
    0.012, the first pass projects it to 0 and the second ends the loop. An
    exit threshold of 0.001 or 0.1 (or >= for <) gives other counts."""
    h = 1.0 / 240
    c = _exit_cloth(oracle_mod, bend=1.0, thickness=THICK, tolerance=TOL, collFlags=CF_WALLS)
    c.add_walls((-10, -1, -10), (10, 10, 10))
    c.init_grid(4, 4, 0.1, (0, 0, 0), True, False)
    x = c.get("pos")
    x[:, 1] = -1.0 - below
    c.put("pos", x)
    c.put("prev", x)
    _cancel_gravity(c, h, 16)
    c.substep(h)
    st = c.stats()
    assert st["iterations"] == iterations
    # the error of the pass that ended the loop
    assert st["contactError"] == pytest.approx(0.008 if iterations == 1 else 0.0, abs=1e-6)


def _folded_pair(theta):
    """Two unit right triangles sharing the edge (0, 1) along z, folded by
    theta about it: triangle (0, 1, 2) in the plane y = 0, vertex 3 rotated."""
    pos = np.array([[0, 0, 0], [0, 0, 1], [1, 0, 0], [-math.cos(theta), math.sin(theta), 0]], float)
    uv = np.array([[0, 0], [0, 1], [1, 0], [-1, 0]], float)
    return pos, uv


@pytest.mark.parametrize("freq,iterations", [(2, 3), (1, 2)])
def test_bend_exit_threshold_and_frequency(oracle_mod, freq, iterations):
    """One bend folded by theta0 = 1.004 rad with s = 0.0005 x 10 = 0.005,
    one end of the hinge pinned (no rigid drift for the FEM pass to undo):
    its first pass reports s theta0 = 0.00502 >= 0.005 and moves the fold by
    s theta0 to theta1 = 0.99898 (error 0.004995 < 0.005). With bend
    frequency 2 the second iteration has no bend pass and keeps the old
    error, so the loop ends after the third; with frequency 1 after the
    second. A threshold of 0.05, or a frequency test of k % f == 1, gives
    other counts."""
    h = 1.0 / 240
    theta0 = 1.004
    c = _exit_cloth(oracle_mod, bend=10.0, bendFreq=freq)
    pos, uv = _folded_pair(theta0)
    c.init_arrays(pos, uv, np.array([0.0, 1.0, 1.0, 1.0]), np.array([[0, 1, 2], [1, 0, 3]], dtype=np.int32))
    assert c.counts()["B"] == 1
    c.put("prev", pos)
    _cancel_gravity(c, h, 4)
    c.substep(h)
    st = c.stats()
    assert st["iterations"] == iterations
    assert st["bendError"] == pytest.approx(0.005 * theta0 * (1 - 0.005), abs=2e-7)
