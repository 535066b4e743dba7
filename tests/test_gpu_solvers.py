"""The reference's Jacobi and conjugate-residual solvers on the GPU
(Physics/ClothPBD.cpp:290-298, 371-478, 616-662): bit-exact against the
oracle frame by frame on a pinned cloth with links (and, ignored by these
solvers as by the reference, FEM triangles, bends and a sphere SDF)."""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["jacobi", "jacobi_cr", "explicit"])
@pytest.mark.parametrize("base", ["drape32", "sdf40"])
def test_solver_bit_exact(sc, oracle_mod, solver, base):
    scene = {"drape32": S.DRAPE32, "sdf40": S.DRAPE256.with_(n=40)}[base]
    # unit 1 and moderate link stiffnesses keep CR finite (it diverges with
    # the defaults, as the reference's does)
    scene = scene.with_(fem=False, unit=1.0, stretch=1.0, young=1.0, bend=1.0)
    print(f"solver {solver} on {base}", flush=True)
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSolver(solver)
    o = oracle_mod.from_scene(scene)
    o.set(solver=m.GetSolver())
    for f in range(10):
        patch.Step(scene.dt)
        o.step(scene.dt)
        gp, op = m.positions(), o.get("pos")
        assert np.isfinite(op).all()
        err = float(np.abs(gp.astype(np.float64) - op).max())
        assert err == 0.0, f"{solver} {base} frame {f}: max abs diff {err:.3e}"


def test_batch_refuses_other_solvers(sc):
    world, patch = S.build_patch(S.DRAPE32, sc)
    patch.GetModel().SetSolver("jacobi")
    b = sc.ClothBatch([patch])
    with pytest.raises(sc.SolidClothError):
        b.Step(S.DRAPE32.dt)
