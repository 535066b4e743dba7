"""The 512x512 self-colliding cloth on the icosphere mesh at full size
(262,144 vertices, 522,242 triangles; obstacle 40,962 vertices, 81,920
triangles): one substep from the GPU's own state after 12 frames against
the oracle bit for bit (broad-phase candidate pairs against the oracle's
grid broad phase, contact sets of every kind, and positions), and the
whole 300-frame drop completing with finite positions, every step taken,
nothing below the ground."""
import time

import numpy as np
import pytest

from solidcloth_amd import scenes as S

pytestmark = pytest.mark.gpu

SCENE = S.MESH512


def _h(scene):
    return float(np.float32(np.float32(scene.dt) / np.float32(scene.substeps)))


def test_mesh512_full_size_substep_parity(sc, oracle_mod):
    from parity_util import same_contacts

    world, patch = S.build_patch(SCENE, sc)
    m = patch.GetModel()
    for f in range(12):
        patch.Step(SCENE.dt)
    print("mesh512: 12 frames on the GPU", flush=True)
    o = oracle_mod.from_scene(SCENE, tri_order=m.constraint_order("tri"), bend_order=m.constraint_order("bend"))
    o.set_contact_order("edge", m.constraint_order("edge_contact"))
    o.set_contact_order("tri", m.constraint_order("tri_contact"))
    o.set(accel=1, numSteps=1, recordCandidates=1)
    p = m.positions().astype(np.float64)
    o.put("pos", p)
    o.put("prev", p)
    o.put("vel", m.velocities().astype(np.float64))
    patch.SetNumSteps(1)
    h = _h(SCENE)
    patch.Step(h)
    t0 = time.perf_counter()
    o.substep_detect(h)
    from test_gpu_self import _same_candidates

    n = _same_candidates(m, o, "mesh512 substep")
    print(f"mesh512: {n} candidate pairs equal as sets", flush=True)
    for kind in ("vt", "ee"):
        ok = o.self_contacts(kind)[0]
        gk = m.self_contacts(kind)[0]
        assert set(ok.tolist()) == set(gk.tolist()), f"{kind} self contact sets differ"
        o.set_self_order(kind, gk)
    o.substep_solve(h)
    print(f"mesh512: oracle substep {time.perf_counter() - t0:.1f} s", flush=True)
    nv, ne, nt = same_contacts(m, o)
    assert nv > 1000 and ne > 1000 and nt > 1000, (nv, ne, nt)
    gp, op = m.positions(), o.get("pos")
    assert np.array_equal(gp.astype(np.float64), op), f"max abs diff {np.abs(gp - op).max():.3e}"


def test_mesh512_300_frames_complete(sc):
    world, patch = S.build_patch(SCENE, sc)
    m = patch.GetModel()
    ground = SCENE.walls_min[1]
    peaks = []
    t0 = time.perf_counter()
    for f in range(1, SCENE.frames + 1):
        patch.Step(SCENE.dt)
        st = m.self_stats()
        peaks.append((st["peak_vt"], st["peak_ee"]))
        if f % 25 == 0:
            p = m.positions()
            print(f"mesh512 frame {f}: {time.perf_counter() - t0:.1f} s, self contacts {peaks[-1]}, "
                  f"vertex contacts {m.GetNumContacts()}, lowest y {p[:, 1].min():.4f}", flush=True)
            assert np.isfinite(p).all(), f
            assert p[:, 1].min() >= ground, f
    v = np.linalg.norm(m.velocities(), axis=1)
    st = m.self_stats()
    print(f"mesh512: 300 frames in {time.perf_counter() - t0:.1f} s, reruns {st['reruns']}, "
          f"largest self contacts {max(p[0] for p in peaks)} / {max(p[1] for p in peaks)}, "
          f"speed max {v.max():.3f}", flush=True)
    assert max(p[1] for p in peaks) > 1000  # the drop does self-collide
    assert v.max() < 5.0  # the cloth has landed: no runaway
