"""The reference's Jacobi and conjugate-residual link solvers in the oracle
(Physics/ClothPBD.cpp:616-662 SolveJacobi / SolveLinksJacobi, :371-478
SolveCR as compiled with SKIP_ALPHA and SKIP_BETA), pinned by their closed
forms on one stretched link: a Jacobi pass moves each end by omega times the
link's projection; the CR passes move the ends along the link by
alpha d m with d = r + beta d, r = -(len - len0) stiffness unit."""
import numpy as np
import pytest


def _link(oracle_mod, solver, iterations, stiffness=50.0, unit=1.0, area=2.0):
    c = oracle_mod.Cloth(real64=True)
    c.set(unit=unit, area=area, fem=0, dihedral=0, stretch=stiffness, shear=0, bend=0, iterations=iterations,
          fixedIterations=1, numSteps=1, solver=solver)
    c.init_grid(1, 2, 1.0, (0, 0, 0), True, False)  # two particles, one link of rest length 1
    pos = c.get("pos").copy()
    pos[1] = pos[0] + np.array([0.0, 0.0, 1.3])
    c.put("pos", pos)
    c.put("prev", pos)
    c.put("vel", np.zeros((2, 3)))
    return c, pos


def test_jacobi_one_pass_closed_form(oracle_mod):
    h, k_s, unit, area = 2.0 ** -7, 50.0, 0.0, 2.0   # unit 0: no gravity, k = 0
    c, x0 = _link(oracle_mod, 0, 1, k_s, 1.0, area)
    c.set(unit=1.0)
    # with gravity on, prediction moves both free particles alike; compare
    # the link's change of length instead of absolute positions
    c.step(h)
    x = c.get("pos")
    # prediction: v = h g, x_pred = x0 + h v (both particles alike)
    g = np.array([0.0, -9.8, 0.0])
    xp = x0 + h * (h * g)
    d = xp[0] - xp[1]
    L = np.linalg.norm(d)
    mu = 1.0 / 2.0
    k = h * h * area * 1.0
    eps = 1.0 / (k * k_s)
    s = 1.0 / (1.0 + mu * eps)
    lam = (L - 1.0) * mu * s
    disp = 0.5 * lam * d / L
    want = xp.copy()
    want[0] -= disp
    want[1] += disp
    assert np.allclose(x, want, atol=1e-9)


def test_cr_two_passes_closed_form(oracle_mod):
    h, k_s = 2.0 ** -7, 3.0
    c, x0 = _link(oracle_mod, 2, 2, k_s)
    c.step(h)
    x = c.get("pos")
    g = np.array([0.0, -9.8, 0.0])
    p = x0 + h * (h * g)
    dd = 0.0
    for it in range(2):
        delta = p[1] - p[0]
        L = np.linalg.norm(delta)
        r = -(L - 1.0) * k_s * 1.0
        n = delta / L
        beta = float(np.float32(np.float32(it) / np.float32(1)) ** np.float32(0.6))
        dd = r + beta * dd
        disp = 0.1 * dd * n
        p = p.copy()
        p[0] -= disp
        p[1] += disp
    assert np.allclose(x, p, atol=1e-9)


@pytest.mark.parametrize("solver", [0, 2, 3, 4])
def test_non_gs_solvers_ignore_triangles_and_contacts(oracle_mod, solver):
    """Jacobi and CR project links only: with FEM triangles, bends and a
    ground plane present the result equals the links-only cloth."""
    def run(fem):
        c = oracle_mod.Cloth(real64=True)
        # moderate link stiffnesses and unit 1: CR's step alpha d grows with
        # both (the defaults drive it to NaN within a frame, as in the reference)
        c.set(unit=1.0, fem=fem, dihedral=1, stretch=1.0, shear=1.0, bend=1.0, iterations=4, fixedIterations=1,
              numSteps=2, solver=solver, collFlags=8)
        c.init_grid(6, 6, 0.1, (0, 0, 0), True, True)
        c.add_walls((-10, -0.05, -10), (10, 10, 10))
        for _ in range(5):
            c.step(1.0 / 60)
        return c.get("pos")
    a, b = run(1), run(0)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
