"""GPU parity of the batched substep: many cloths of uneven size stepped by
one launch per substep (one workgroup per cloth), through the C ABI.

Each cloth of a batch must end where the oracle ends when the oracle is run
with that cloth's constraints in the GPU's colour order, and where the same
patch ends when stepped alone with the colour sweep (mode 0): bit-exact, on
top of the 1e-5-of-extent bar. Every launch configuration (block size,
positions staged in LDS or not) gives the same bits, which is what lets the
autotuner pick by time alone.
"""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

from parity_util import same_positions

pytestmark = pytest.mark.gpu


def _build(sc, scenes):
    built = [S.build_patch(s, sc) for s in scenes]
    return built, sc.ClothBatch([p for _, p in built])


def test_batch_matches_oracle_100_substeps(sc, oracle_mod):
    """Six small cloths (two of them with the reference's early exit) against
    the oracle for 25 frames (100 substeps)."""
    scenes = S.batch_scenes(6, seed=11, n_range=(16, 24), frames=25)
    scenes = [s.with_(fixed_iterations=(k % 3 != 0)) for k, s in enumerate(scenes)]
    built, batch = _build(sc, scenes)
    batch.SetLaunch(256, 1)
    batch.Step(scenes[0].dt)  # selects the colour sweep on every patch
    oracles = []
    for s, (_, p) in zip(scenes, built):
        m = p.GetModel()
        o = oracle_mod.from_scene(s, tri_order=m.constraint_order("tri"), bend_order=m.constraint_order("bend"))
        o.step(s.dt)
        oracles.append(o)
    for f in range(1, 25):
        batch.Step(scenes[0].dt)
        for o in oracles:
            o.step(scenes[0].dt)
    for k, (s, (_, p), o) in enumerate(zip(scenes, built, oracles)):
        m = p.GetModel()
        same_positions(m, o, f"cloth {k} (n={s.n}) after 100 substeps")
        assert np.array_equal(m.velocities().astype(np.float64), o.get("vel"))
        assert m.stats()["iterations"] == o.stats()["iterations"]
    # the pins held
    for s, (_, p) in zip(scenes, built):
        pos = p.GetModel().positions()
        ref = S.build_patch(s, sc)[1].GetModel().positions()
        assert np.array_equal(pos[list(s.pins)], ref[list(s.pins)])


def test_batch_equals_single_patch_colour_sweep(sc):
    """A batch of 24 cloths (sides 16-128, so some exceed the LDS and stay in
    HBM) against the same cloths stepped one by one with the colour sweep."""
    scenes = S.batch_scenes(24, seed=5, n_range=(16, 128))
    built, batch = _build(sc, scenes)
    batch.SetLaunch(512, 1)
    singles = [S.build_patch(s, sc) for s in scenes]
    for _, p in singles:
        p.GetModel().SetSweep(0)
    for f in range(6):
        batch.Step(scenes[0].dt)
        for _, p in singles:
            p.Step(scenes[0].dt)
    for k, ((_, pb), (_, ps)) in enumerate(zip(built, singles)):
        a, b = pb.GetModel().positions(), ps.GetModel().positions()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), f"cloth {k} (n={scenes[k].n}): max diff {np.abs(a - b).max():.3e}"


def test_every_launch_configuration_gives_the_same_bits(sc):
    scenes = S.batch_scenes(16, seed=7, n_range=(16, 128))
    results = {}
    for bs in (256, 512, 1024):
        for stage in (0, 1):
            built, batch = _build(sc, scenes)
            batch.SetLaunch(bs, stage)
            for f in range(3):
                batch.Step(scenes[0].dt)
            results[(bs, stage)] = np.concatenate([p.GetModel().positions() for _, p in built])
    first = results[(256, 0)]
    assert np.isfinite(first).all()
    for key, r in results.items():
        assert np.array_equal(r, first), f"launch {key} differs"


def test_autotuner_times_every_candidate_and_keeps_results(sc):
    scenes = S.batch_scenes(16, seed=9, n_range=(16, 96))
    built, batch = _build(sc, scenes)
    ref_built, ref = _build(sc, scenes)
    ref.SetLaunch(256, 0)
    for f in range(8):
        batch.Step(scenes[0].dt)
        ref.Step(scenes[0].dt)
    bs, stage, tuned = batch.GetLaunch()
    assert tuned and bs in (256, 512, 1024) and stage in (0, 1)
    t = batch.tuning()
    assert sorted((b, s) for b, s, _ in t) == [(b, s) for b in (256, 512, 1024) for s in (0, 1)]
    assert min(ms for _, _, ms in t) == [ms for b, s, ms in t if (b, s) == (bs, stage)][0]
    for (_, p), (_, q) in zip(built, ref_built):
        assert np.array_equal(p.GetModel().positions(), q.GetModel().positions())


def test_batch_workgroup_times(sc):
    """Per-cloth workgroup clocks: one entry per patch, every workgroup ends
    after it starts, and recording leaves the results unchanged (the same
    cloths stepped without it give the same bits)."""
    scenes = S.batch_scenes(12, seed=3, n_range=(16, 48))
    built, batch = _build(sc, scenes)
    built2, batch2 = _build(sc, scenes)
    for b in (batch, batch2):
        b.SetLaunch(256, 1)
    assert batch.workgroup_times() is None
    batch.RecordWorkgroupTimes(True)
    for _ in range(3):
        batch.Step(scenes[0].dt)
        batch2.Step(scenes[0].dt)
    wt = batch.workgroup_times()
    assert wt["start_us"].shape == (12,) and wt["unit"].shape == (12, 4)
    assert (wt["start_us"] >= 0).all() and (wt["end_us"] > wt["start_us"]).all()
    assert wt["start_us"].min() == 0.0
    for (_, p), (_, q) in zip(built, built2):
        assert np.array_equal(p.GetModel().positions(), q.GetModel().positions())
