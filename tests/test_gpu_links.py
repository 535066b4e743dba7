"""GPU parity of the link (non-FEM) stretch model (Physics/ClothPBD.cpp:9-50,
the reference's default model): the oracle runs its links in the GPU's
Gauss-Seidel order (`scGetConstraintOrder` kind 4) and positions,
velocities, iteration counts and contact sets must be bit-exact (the 1e-5
bar of the cloth extent is asserted as well). Both sweep schedules, with
dihedral bends and with the grid's bend links."""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

from parity_util import same_contacts, same_positions

pytestmark = pytest.mark.gpu


def _pair(sc, oracle_mod, scene, mode):
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(mode)
    o = oracle_mod.from_scene(scene, bend_order=m.constraint_order("bend"))
    o.set_order("link", m.constraint_order("link"))
    return world, patch, m, o


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("dihedral", [True, False])
def test_link_model_100_substeps_bit_exact(sc, oracle_mod, mode, dihedral):
    scene = S.DRAPE32.with_(fem=False, dihedral=dihedral, stretch=1e5)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for _ in range(25):
        patch.Step(scene.dt)
        o.step(scene.dt)
    same_positions(m, o, f"links (dihedral={dihedral}, mode {mode}) after 100 substeps")
    assert np.array_equal(m.velocities().astype(np.float64), o.get("vel"))
    assert m.stats()["linkError"] == np.float32(o.stats()["linkError"])
    assert m.stats()["linkError"] > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_link_model_reference_early_exit(sc, oracle_mod, mode):
    """The reference's exit test with the link error (cl < 0.05)."""
    scene = S.DRAPE32.with_(fem=False, fixed_iterations=False)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for _ in range(10):
        patch.Step(scene.dt)
        o.step(scene.dt)
        assert m.stats()["iterations"] == o.stats()["iterations"]
    same_positions(m, o, "links, reference exit")


def test_link_model_with_contacts(sc, oracle_mod):
    scene = S.DRAPE256.with_(n=40, fem=False, frames=40)
    world, patch, m, o = _pair(sc, oracle_mod, scene, 1)
    for f in range(40):
        patch.Step(scene.dt)
        o.step(scene.dt)
    same_positions(m, o, "links over sphere SDF and ground")
    nv, _, _ = same_contacts(m, o)
    assert nv > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_error_statistics_are_the_last_passes(sc, oracle_mod, mode):
    """With a fixed iteration count the reported bend, contact and link
    errors are those of the last pass of each kind (ClothPBD.cpp:337-369),
    also when the last iteration's contacts are applied by the velocity
    kernel (slab sweep, mode 1)."""
    for fem in (True, False):
        scene = S.DRAPE256.with_(n=40, fem=fem)
        world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
        if fem:
            o = oracle_mod.from_scene(scene, tri_order=m.constraint_order("tri"),
                                      bend_order=m.constraint_order("bend"))
        for _ in range(30):
            patch.Step(scene.dt)
            o.step(scene.dt)
        same_positions(m, o, f"fem={fem}, mode {mode}")
        gs, os_ = m.stats(), o.stats()
        for key in ("bendError", "contactError", "linkError"):
            assert gs[key] == np.float32(os_[key]), (key, gs[key], os_[key])
        assert gs["contactError"] > 0 and gs["bendError"] > 0
