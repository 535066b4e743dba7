"""The mouse spring (Physics/ClothPBD.cpp:267-282, ClothModel.inl:86-98) in
the oracle, pinned by its closed form: right after the prediction the
spring's three particles move by -h a_k 100 (|p - m|) (p - m)/|p - m|, where p
is the barycentric point of their positions and m the mouse point; pinned
particles move too. Gravity is switched off (unit 0) and every constraint has
zero stiffness, so the sweeps change nothing."""
import numpy as np
import pytest


def _cloth(oracle_mod, attached):
    c = oracle_mod.Cloth(real64=True)
    c.set(unit=0.0, fem=0, dihedral=0, stretch=0.0, shear=0.0, bend=0.0, iterations=1, fixedIterations=1,
          numSteps=1)
    c.init_grid(3, 3, 0.5, (0.0, 0.0, 0.0), True, attached)
    return c


@pytest.mark.parametrize("attached", [False, True])
def test_mouse_spring_closed_form(oracle_mod, attached):
    c = _cloth(oracle_mod, attached)
    tri, a, m = (6, 7, 3), (0.2, 0.3, 0.5), np.array([0.4, 0.7, -0.2])
    x0 = c.get("pos")
    c.set_mouse_spring(tri, a, m)
    h = 1.0 / 64
    c.step(h)
    x = c.get("pos")
    # the spring keeps its weights and point in FP32 (ClothTypes.h:85-91)
    a = np.float32(a).astype(np.float64)
    m = np.float32(m).astype(np.float64)
    p = a[0] * x0[tri[0]] + a[1] * x0[tri[1]] + a[2] * x0[tri[2]]
    d = p - m
    L = np.linalg.norm(d)
    disp = 100.0 * L * (d / L)
    want = x0.copy()
    for k in range(3):
        want[tri[k]] -= h * a[k] * disp
    assert np.allclose(x, want, rtol=0, atol=1e-12)
    moved = set(np.nonzero(np.abs(x - x0).max(axis=1) > 0)[0].tolist())
    assert moved == set(tri)
    # removing it leaves the cloth at rest from then on
    c.remove_mouse_spring()
    c.put("vel", np.zeros_like(x))
    before = c.get("pos")
    c.step(h)
    assert np.array_equal(c.get("pos"), before)


def test_mouse_spring_pulls_towards_point(oracle_mod):
    """Repeated substeps bring the barycentric point closer to the mouse."""
    c = _cloth(oracle_mod, False)
    tri, a, m = (0, 1, 3), (1 / 3, 1 / 3, 1 / 3), np.array([0.0, 1.0, 0.0])
    c.set_mouse_spring(tri, a, m)
    dist = []
    for _ in range(5):
        x = c.get("pos")
        dist.append(np.linalg.norm(sum(a[k] * x[tri[k]] for k in range(3)) - m))
        c.step(1.0 / 240)
    assert all(d2 < d1 for d1, d2 in zip(dist, dist[1:]))
