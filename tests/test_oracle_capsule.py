"""The capsule collidable in the oracle (Physics/ClothCollisionSDF.h:57-122):
its closest-point queries pinned bit for bit to the reference's compiled
Geometry code (tests/golden/ref_segm_triangle.npz, written by
tests/golden/make_golden.py), its axis from the rotation quaternion pinned by
closed forms, and a vertex contact against it by its closed form."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_segm_triangle.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def test_segm_triangle_matches_reference(oracle_mod, golden):
    P, want = golden["points"], golden["segm_tri_out"]
    for k in range(len(P)):
        d, t, bar, ce, ct = oracle_mod.closest_pt_segm_triangle(*P[k])
        got = np.concatenate([[d, t], bar, ce, ct]).astype(np.float32)
        assert np.array_equal(got, want[k], equal_nan=True), (k, got, want[k])


def test_point_segm_matches_reference(oracle_mod, golden):
    P, want = golden["points"], golden["point_segm_out"]
    for k in range(len(P)):
        t, cp = oracle_mod.closest_pt_segm(P[k, 4], P[k, 0], P[k, 1])
        got = np.concatenate([[t], cp]).astype(np.float32)
        assert np.array_equal(got, want[k], equal_nan=True), k


@pytest.mark.parametrize("rot,axis", [
    ((1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0)),                         # identity: the y axis
    ((np.cos(np.pi / 4), np.sin(np.pi / 4), 0.0, 0.0), (0.0, 0.0, 1.0)),   # 90 deg about x: y -> z
    ((np.cos(np.pi / 4), 0.0, 0.0, np.sin(np.pi / 4)), (-1.0, 0.0, 0.0)),  # 90 deg about z: y -> -x
    ((np.cos(np.pi / 6), 0.0, 0.0, np.sin(np.pi / 6)), (-np.sin(np.pi / 3), np.cos(np.pi / 3), 0.0)),
])
def test_capsule_axis_from_rotation(oracle_mod, rot, axis):
    c = oracle_mod.Cloth()
    center, hh = np.array([0.1, 0.2, -0.3]), 0.25
    _, ends = c.add_capsule(center, 0.1, hh, rot)
    want_p = center + hh * np.array(axis)
    want_q = center - hh * np.array(axis)
    assert np.allclose(ends[0], want_p, atol=2e-7) and np.allclose(ends[1], want_q, atol=2e-7)
    if rot == (1.0, 0.0, 0.0, 0.0):  # no rounding at all without rotation
        assert np.array_equal(ends[0], np.float32(center + [0, hh, 0]))


def test_capsule_vertex_contact_closed_form(oracle_mod):
    """A horizontal cloth just above a capsule lying along x: every vertex
    within the contact radius gets the surface point axis + r n, with n the
    unit vector from its nearest axis point."""
    c = oracle_mod.Cloth(real64=True)
    r, thick, tol = 0.1, 0.01, 0.02
    c.set(unit=0.0, fem=0, dihedral=0, stretch=0.0, shear=0.0, bend=0.0, iterations=1, fixedIterations=1,
          numSteps=1, collFlags=1, thickness=thick, tolerance=tol)
    c.init_grid(5, 5, 0.05, (0.0, r + 0.012, 0.0), True, False)
    rot = (np.cos(np.pi / 4), 0.0, 0.0, np.sin(np.pi / 4))  # axis along x
    _, ends = c.add_capsule((0.0, 0.0, 0.0), r, 0.3, rot)
    c.substep_detect(1.0 / 240)
    ids, data = c.contacts()
    pos = c.get("prev")
    assert len(ids) > 0
    for (v, _), d in zip(ids, data):
        x = pos[v]
        a = np.array([x[0], 0.0, 0.0])  # nearest axis point (|x| < hh)
        n = (x - a) / np.linalg.norm(x - a)
        assert np.allclose(d[3:6], n, atol=1e-6), (v, d[3:6], n)
        assert np.allclose(d[0:3], a + r * n, atol=1e-6)
    dist = np.hypot(pos[:, 1], pos[:, 2]) - r
    want = set(np.nonzero((dist >= 0) & (dist < thick + tol))[0].tolist())
    assert set(int(v) for v, _ in ids) == want
