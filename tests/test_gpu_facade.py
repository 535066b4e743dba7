"""The C++ classes of include/solidcloth.hpp (ClothPatch, ClothModel,
ClothModelPBD, CollisionWorld with Sphere and Walls descriptions) step a
drape through the library exactly as the Python binding does: the program
tests/cpp/facade_drape (built by __graft_entry__.build()) and the binding run
the same scene on the GPU, and the final positions agree bit for bit. The
program also checks its getters against its setters, that a refused call
throws solidcloth::Error, and steps its last frame as four single substeps of
ClothModelPBD::Step."""
import os
import subprocess

import numpy as np
import pytest

from solidcloth_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "facade_drape")


def _f(x):
    # the float32 value, printed exactly (strtof gives it back unchanged)
    return repr(float(np.float32(x)))


@pytest.mark.gpu
def test_facade_drape_matches_the_binding(sc, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/facade_drape is not built: run __graft_entry__.build()")
    scene = S.DRAPE256.with_(n=48, analytic_sphere=True, sdf_cells=0, fixed_iterations=False, frames=30)
    out = tmp_path / "facade_pos.bin"
    args = [EXE, str(scene.n), _f(scene.inc), _f(scene.offset[1]), _f(scene.young), _f(scene.poisson),
            _f(scene.bend), str(scene.iterations), _f(scene.thickness), _f(scene.tolerance), _f(scene.friction),
            _f(scene.sphere_radius), _f(scene.walls_min[1]), str(scene.frames), str(out)]
    print("facade_drape", " ".join(args[1:-1]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refused:" in r.stdout
    got = np.fromfile(out, dtype=np.float32).reshape(-1, 3)

    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    for _ in range(scene.frames):
        patch.Step(scene.dt)
    want = m.positions()
    assert got.shape == want.shape
    assert m.GetNumContacts() > 0  # the cloth has reached the sphere
    assert np.array_equal(got, want), f"max |d| = {np.abs(got - want).max():.3e}"
