"""GPU parity of edge and triangle contacts against analytic spheres and grid
SDFs (Physics/ClothCollision.cpp:472-579): sphere segment / triangle queries,
Frank-Wolfe triangle queries on the grid SDF, the border filters, and their
Gauss-Seidel passes, against the CPU oracle through the C ABI."""
import pytest

from solidcloth_amd import scenes as S

from parity_util import pair, same_contacts, same_positions

pytestmark = pytest.mark.gpu
V, T, E, W = S.CF_VERTICES, S.CF_TRIANGLES, S.CF_EDGES, S.CF_WALLS


@pytest.mark.parametrize("name,scene", [
    ("sdf-VTEW", S.DRAPE256.with_(n=40, flags=V | T | E | W)),
    ("sphere-VTEW", S.DRAPE256.with_(n=40, flags=V | T | E | W, analytic_sphere=True)),
    ("sphere-TEW", S.DRAPE256.with_(n=40, flags=T | E | W, analytic_sphere=True)),
], ids=lambda x: x if isinstance(x, str) else "")
def test_sdf_and_sphere_feature_contacts_bit_exact(sc, oracle_mod, name, scene):
    print(f"{name}", flush=True)
    world, patch, m, o = pair(sc, oracle_mod, scene)
    seen = [0, 0, 0]
    for f in range(40):
        patch.Step(scene.dt)
        o.step(scene.dt)
        seen = [a + b for a, b in zip(seen, same_contacts(m, o))]
        same_positions(m, o, f"{name} frame {f}")
    assert seen[2] > 0 and (seen[1] > 0 or "sdf" in name), seen

