"""A mesh-based patch (ClothPatch::Init(const Mesh&)) stepped on the GPU:
bit-exact against the oracle built from the same mesh and run in the GPU's
constraint order, and its render mesh keeps the unwelded vertices (the seam's
copies follow their particle) with angle-weighted normals within 1e-5 of the
oracle's."""
import numpy as np
import pytest

from mesh_patch_util import seamed_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sweep", [0, 1])
def test_mesh_patch_bit_exact(sc, oracle_mod, sweep):
    verts, tris, uvs, colors = seamed_grid(17)
    print(f"mesh patch sweep {sweep}", flush=True)
    world = sc.CollisionWorld()
    patch = sc.ClothPatch(world)
    m = patch.GetModel()
    for name, v in (("SetUnit", 1.0), ("SetShearStiffness", 1e4), ("SetPoisson", 0.3), ("SetBendStiffness", 0.01),
                    ("SetNumIterations", 10)):
        getattr(m, name)(v)
    m.SetFEM(True)
    m.SetDihedral(True)
    m.SetFixedIterations(True)
    patch.SetNumSteps(4)
    mesh = sc.Mesh(verts, tris)
    mesh.set_uvs(uvs)
    mesh.set_colors(colors)
    patch.InitFromMesh(mesh, (0.0, 0.0, 0.0), True)
    m.SetSweep(sweep)
    o = oracle_mod.Cloth()
    o.set(unit=1.0, shear=1e4, poisson=0.3, bend=0.01, iterations=10, fem=1, dihedral=1, fixedIterations=1,
          numSteps=4)
    mp = o.init_mesh(verts, tris, uvs, colors, (0.0, 0.0, 0.0), True)
    o.set_order("tri", m.constraint_order("tri"))
    o.set_order("bend", m.constraint_order("bend"))
    dt = 1.0 / 60
    for f in range(15):
        patch.Step(dt)
        o.step(dt)
        err = float(np.abs(m.positions().astype(np.float64) - o.get("pos")).max())
        assert err == 0.0, f"frame {f}: max abs diff {err:.3e}"
    patch.UpdateMesh()
    rv, rn, ri = patch.mesh_arrays()
    assert len(rv) == len(verts)
    assert np.array_equal(rv, m.positions()[mp])
    ref = oracle_mod.mesh_normals(rv, ri)
    assert float(np.abs(rn - ref).max()) <= 1e-5
