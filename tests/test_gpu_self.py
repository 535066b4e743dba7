"""GPU parity of continuous self collision (LBVH refit + overlap walks, FP64
CCD cubic, Jones-Plassmann colouring of the contacts, coloured projections)
against the CPU oracle's exhaustive candidate enumeration, substep by
substep: the broad-phase candidate pairs must be the oracle's closed-interval
overlap set exactly; the GPU solves its contacts colour group by colour group
and hands that order to the oracle, which must then produce the same contacts
(keys, ids, weights, normals, sides, lambdas) and the same positions bit for
bit."""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

from parity_util import pair, same_positions

pytestmark = pytest.mark.gpu


def _rows(keys, ids, data):
    return sorted((int(k), tuple(int(x) for x in i), tuple(np.float32(d).tolist())) for k, i, d in zip(keys, ids, data))


def _same_candidates(m, o, what):
    """The GPU's broad-phase candidate pairs (LBVH overlap walks) equal the
    oracle's exact closed-interval overlap set, as sets, for both kinds."""
    n = 0
    for kind in ("vt", "ee"):
        g = m.self_candidates(kind)
        r = o.self_candidates(kind)
        r = r[np.lexsort((r[:, 1], r[:, 0]))] if len(r) else r.reshape(0, 2)
        assert len(np.unique(g, axis=0)) == len(g), f"{what}: {kind} pair listed twice"
        assert np.array_equal(g, r), f"{what}: {kind} candidate sets differ ({len(g)} vs {len(r)})"
        n += len(g)
    return n


def _run(sc, oracle_mod, scene, mode, substeps):
    world, patch, m, o = pair(sc, oracle_mod, scene, mode)
    o.set(recordCandidates=1)
    patch.SetNumSteps(1)
    h = float(np.float32(np.float32(scene.dt) / np.float32(scene.substeps)))
    seen = [0, 0]
    cands = 0
    for k in range(substeps):
        patch.Step(h)
        o.substep_detect(h)
        cands += _same_candidates(m, o, f"substep {k}")
        for kind, j in (("vt", 0), ("ee", 1)):
            o.set_self_order(kind, m.self_contacts(kind)[0])
        o.substep_solve(h)
        for kind, j in (("vt", 0), ("ee", 1)):
            gk, gi, gd = m.self_contacts(kind)
            ok, oi, od = o.self_contacts(kind)
            assert _rows(gk, gi, gd) == _rows(ok, oi, od), f"{kind} contacts differ at substep {k}"
            # the GPU's list is its processing order, the oracle's too
            assert np.array_equal(gk, ok), f"{kind} order at substep {k}"
            seen[j] += len(gk)
        same_positions(m, o, f"substep {k}")
    assert cands > 0
    return seen, m, o


@pytest.mark.parametrize("mode", [0, 1])
def test_fold_self_collision_bit_exact(sc, oracle_mod, mode):
    print(f"fold24 mode={mode}", flush=True)
    seen, m, o = _run(sc, oracle_mod, S.FOLD24, mode, 48)
    assert seen[0] > 0 and seen[1] > 0, seen


def test_fold_self_collision_own_edge_weights_bit_exact(sc, oracle_mod):
    """The same fold with each self edge-edge end moved by its own weight
    (scSetSelfEdgeWeights 1)."""
    print("fold24 own edge weights", flush=True)
    seen, m, o = _run(sc, oracle_mod, S.FOLD24.with_(self_edge_weights=1), 1, 48)
    assert seen[1] > 0, seen


def test_fold_self_collision_reference_early_exit(sc, oracle_mod):
    """The exit test includes the self contact errors against the threshold."""
    scene = S.FOLD24.with_(fixed_iterations=False, iterations=15)
    world, patch, m, o = pair(sc, oracle_mod, scene)
    patch.SetNumSteps(1)
    h = float(np.float32(np.float32(scene.dt) / np.float32(scene.substeps)))
    for k in range(48):
        patch.Step(h)
        o.substep_detect(h)
        o.set_self_order("vt", m.self_contacts("vt")[0])
        o.set_self_order("ee", m.self_contacts("ee")[0])
        o.substep_solve(h)
        assert m.stats()["iterations"] == o.stats()["iterations"], k
    same_positions(m, o, "fold early exit")


def test_mesh_obstacle_with_self_collision_mid_size(sc, oracle_mod):
    """The mesh512 workload's path at 40x40: perturbed cloth on an icosphere
    mesh (subdivision 2) with vertex, edge, triangle and self contacts; self
    contacts begin after about 80 substeps, once the cloth folds on the
    ground."""
    scene = S.MESH512.with_(n=40, side=1.0, mesh_subdiv=2, mesh_radius=0.25, offset=(0.0, 0.27, 0.0),
                            walls_min=(-10.0, -0.3, -10.0), thickness=0.004, tolerance=0.005, frames=24,
                            perturb_amplitude=0.02, ccd=S.ccd_for_spacing(1.0 / 39), drag=1.0,
                            self_edge_weights=0)
    print("mesh+self 40", flush=True)
    seen, m, o = _run(sc, oracle_mod, scene, 1, scene.frames * scene.substeps)
    assert sum(seen) > 0, seen


def test_self_overflow_reruns_the_step(sc):
    """A step whose self contacts outgrow the contact buffers is rerun from
    its starting state with the buffers grown: a patch that starts with room
    for 8 contacts per kind ends bit-identical to one with the automatic
    capacity, reports the reruns, and never exposes a half-advanced state."""
    scene = S.FOLD24
    runs = []
    for cap in (0, 8):
        print(f"fold24 overflow rerun, capacity {cap}", flush=True)
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        m.SetSelfCapacity(cap)
        peaks = []
        for f in range(scene.frames):
            patch.Step(scene.dt)
            st = m.self_stats()
            peaks.append((st["peak_vt"], st["peak_ee"]))
        runs.append((m.positions(), m.velocities(), m.self_stats(), peaks))
    (p0, v0, s0, k0), (p1, v1, s1, k1) = runs
    assert max(max(k) for k in k0) > 8, "the scene must find more than 8 contacts of a kind in a substep"
    assert s0["reruns"] == 0, s0
    assert s1["reruns"] >= 1 and s1["capacity"] > 8, s1
    assert k0 == k1, "both runs must see the same contact peaks step by step"
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1)
