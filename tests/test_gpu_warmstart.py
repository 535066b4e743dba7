"""GPU checks of the learned warm-start through the C ABI: the BF16-MFMA
forward and backward kernels and Adam against the NumPy oracle, the features
and the FP64 residual against the oracle, the plain-solver fallback, and what
a trained network gains over the method without it.

Tolerances: the GPU rounds every product operand to BF16 and accumulates in
FP32; the oracle with `emulate=True` rounds at the same points and
accumulates exactly, so the two differ by FP32 accumulation and by BF16
rounding flips of values within an FP32 rounding of a tie (bound 5e-3 of the
output scale, 1e-2 in gradient norm). Against the unrounded FP64 network the
bound is the BF16 one (5e-2).
"""
import numpy as np
import pytest

from oracle import warmstart as W
from solidcloth_amd import scenes as S

pytestmark = pytest.mark.gpu


def _data(n, seed=0, scale=0.25):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, W.NUM_PARAMS).astype(np.float32)
    X = rng.normal(0.0, 1.0, (n, 12)).astype(np.float32)
    T = rng.normal(0.0, 0.5, (n, 3)).astype(np.float32)
    w = (rng.random(n) > 0.25).astype(np.float32)
    return p, X, T, w


def test_forward_matches_oracle(sc):
    p, X, _, _ = _data(4099, seed=1)
    ws = sc.WarmStart(1)
    ws.set_params(p)
    assert np.array_equal(ws.params(), p)
    for n in (1, 37, 1000, 4099):
        y = ws.forward(X[:n])
        ye, _ = W.forward(p, X[:n], emulate=True)
        y64, _ = W.forward(p, X[:n], emulate=False)
        scale = float(np.abs(y64).max())
        assert np.isfinite(y).all()
        assert np.abs(y - ye).max() <= 5e-3 * scale, n
        assert np.abs(y - y64).max() <= 5e-2 * scale, n


def test_gradient_matches_oracle(sc):
    p, X, T, w = _data(3001, seed=2)
    ws = sc.WarmStart(2)
    ws.set_params(p)
    loss, g = ws.gradient(X, T, w)
    le, ge = W.loss_and_grad(p, X, T, w, emulate=True)
    l64, g64 = W.loss_and_grad(p, X, T, w, emulate=False)
    assert loss == pytest.approx(le, rel=1e-3)
    assert np.linalg.norm(g - ge) <= 1e-2 * np.linalg.norm(ge)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)
    # every parameter block is covered (weights and biases of all layers)
    for a, b in ((0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 5184), (5184, 5187)):
        assert np.linalg.norm(g[a:b] - ge[a:b]) <= 2e-2 * np.linalg.norm(ge[a:b]) + 1e-7, (a, b)


def test_adam_step_matches_oracle(sc):
    p, X, T, w = _data(2000, seed=4)
    ws = sc.WarmStart(4)
    ws.set_params(p)
    ws.SetLearningRate(1e-3)
    assert ws.steps() == 0
    ws.train(X, T, w, steps=1)
    assert ws.steps() == 1
    p1 = ws.params()
    _, ge = W.loss_and_grad(p, X, T, w, emulate=True)
    pe, _, _ = W.adam(p.astype(np.float64), ge, np.zeros_like(ge), np.zeros_like(ge), 1, 1e-3)
    big = np.abs(ge) > 1e-3 * np.abs(ge).max()
    np.testing.assert_allclose(p1[big], pe[big], atol=2e-6)


def test_training_reduces_the_loss(sc):
    """Targets from a fixed teacher network: 300 Adam steps on the GPU cut the
    loss well below the start."""
    tp, X, _, _ = _data(4096, seed=5, scale=0.2)
    T = W.forward(tp, X, emulate=False)[0].astype(np.float32)
    ws = sc.WarmStart(6)
    l0, _ = ws.gradient(X, T)
    ws.SetLearningRate(3e-3)
    ws.train(X, T, steps=300)
    l1, _ = ws.gradient(X, T)
    assert l1 < 0.2 * l0, (l0, l1)


def _small(n=32, **kw):
    return S.DRAPE256.with_(n=n, **kw)


def test_features_and_scale_match_oracle(sc, oracle_mod):
    scene = _small()
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    ws = sc.WarmStart(0)
    m.SetWarmStart(ws, "train")
    for _ in range(20):  # until the cloth lies on the sphere
        patch.Step(scene.dt)
    prev = m.positions()
    corr = m.warm_data("corr")
    patch.SetNumSteps(1)
    patch.Step(scene.dt / scene.substeps)
    X = m.warm_data("features")
    xhat = m.warm_data("xhat")
    ell, nfree = m.warm_data("scale")
    n = scene.n
    d = float(scene.inc)
    e_axis, e_diag = 2 * n * (n - 1), (n - 1) ** 2
    assert ell == pytest.approx(d * (e_axis + np.sqrt(2) * e_diag) / (e_axis + e_diag), rel=1e-6)
    assert nfree == scene.num_vertices  # drape: nothing pinned
    o = oracle_mod.from_scene(scene)
    idx, rest = o.triangles()
    Xo = W.features(xhat, prev, corr, idx, rest, W.one_ring(len(xhat), o.edges()), ell)
    Xb = W.bf16(Xo.astype(np.float32))
    assert np.abs(Xo[:, 3:12]).max() > 0.05  # corrections and strains are not trivial
    np.testing.assert_allclose(X, Xb, rtol=2.0 ** -7, atol=2e-5)
    # targets of the substep: the correction in output units, weight 1
    Tg = m.warm_data("targets")
    np.testing.assert_allclose(Tg[:, :3], W.SCALE_OUT * (m.positions() - xhat) / ell, rtol=1e-5, atol=1e-6)
    assert np.all(Tg[:, 3] == 1.0)


def test_residual_matches_oracle(sc, oracle_mod):
    scene = _small()
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    for _ in range(5):
        patch.Step(scene.dt)
    o = oracle_mod.from_scene(scene)
    idx, rest = o.triangles()
    s = float(np.float32(0.0005) * np.float32(scene.bend))
    r = m.residual()
    ro = W.residual(m.positions(), idx, rest, o.bends(), s)
    assert ro > 0
    assert r == pytest.approx(ro, rel=1e-9)
    np.testing.assert_allclose(m.residual_terms(), W.residual_terms(m.positions(), idx, rest, o.bends(), s),
                               rtol=1e-9)


def test_train_mode_leaves_the_plain_trajectory(sc):
    scene = _small()
    _, a = S.build_patch(scene, sc)
    _, b = S.build_patch(scene, sc)
    ws = sc.WarmStart(0)
    b.GetModel().SetWarmStart(ws, "train")
    for _ in range(6):
        a.Step(scene.dt)
        b.Step(scene.dt)
    assert np.array_equal(a.GetModel().positions(), b.GetModel().positions())
    assert ws.steps() == 6 * scene.substeps


def test_fallback_reruns_the_plain_step_bit_for_bit(sc):
    scene = _small()
    _, plain = S.build_patch(scene, sc)
    _, judged = S.build_patch(scene, sc)
    _, free = S.build_patch(scene, sc)
    ws = sc.WarmStart(0)
    mj, mf = judged.GetModel(), free.GetModel()
    mj.SetWarmStart(ws, "solve", 4)
    mj.SetWarmStartFallback(0.0)  # every warm-started step is rejected
    mf.SetWarmStart(ws, "solve", 4)
    mf.SetWarmStartFallback(1e9)  # none is
    for _ in range(5):
        plain.Step(scene.dt)
        judged.Step(scene.dt)
        free.Step(scene.dt)
    assert np.array_equal(plain.GetModel().positions(), mj.positions())
    assert np.array_equal(plain.GetModel().velocities(), mj.velocities())
    st = mj.warm_stats()
    assert st["fallbacks"] == 5 and st["last_plain"]
    assert st["residual"] == pytest.approx(plain.GetModel().residual(), rel=1e-12)
    sf = mf.warm_stats()
    assert sf["fallbacks"] == 0 and not sf["last_plain"]
    assert not np.array_equal(plain.GetModel().positions(), mf.positions())
    assert np.isfinite(mf.positions()).all()


def test_plain_judge_keeps_the_better_step(sc):
    """The default judge (scSetWarmStartFallback < 0) runs the plain step from
    the same start and keeps the warm-started result only when its residual
    is not larger: step by step, from one saved state, the judged step's
    positions equal the plain step's or the unjudged warm-started step's,
    as their residuals decide."""
    scene = _small()
    ws = sc.WarmStart(0)  # untrained: the plain solver wins some steps
    _, p = S.build_patch(scene, sc)
    m = p.GetModel()
    m.SetWarmStart(ws, "track")
    for _ in range(3):
        p.Step(scene.dt)
    kept = fell = 0
    for step in range(6):
        m.save_state()
        m.SetWarmStart(ws, "track")
        p.Step(scene.dt)
        pp, rp = m.positions(), m.residual()
        m.restore_state()
        m.SetWarmStart(ws, "solve", 4)
        m.SetWarmStartFallback(float("inf"))  # no judge: the warm-started step as it comes
        p.Step(scene.dt)
        pw, rw = m.positions(), m.residual()
        m.restore_state()
        m.SetWarmStartFallback(-1.0)
        p.Step(scene.dt)
        st = m.warm_stats()
        assert st["steps"] == 1
        assert st["plain_residual"] == pytest.approx(rp, rel=1e-12)
        assert st["warm_residual"] == pytest.approx(rw, rel=1e-12)
        if rw <= rp:
            kept += 1
            assert not st["last_plain"] and st["fallbacks"] == 0 and np.array_equal(m.positions(), pw), step
        else:
            fell += 1
            assert st["last_plain"] and st["fallbacks"] == 1 and np.array_equal(m.positions(), pp), step
    assert kept + fell == 6


def _residual_after(patch, dt, setup, out):
    m = patch.GetModel()
    m.restore_state()
    setup(m)
    patch.Step(dt)
    out.append(m.positions().astype(np.float64))
    return m.residual()


def test_trained_network_gains_over_the_plain_and_untrained(sc):
    """Trained on drape variants (other sphere radius, height and stiffness),
    evaluated on a held-out drape: from the same saved state, 4 warm-started
    iterations leave a smaller residual than 4 plain iterations, and than 4
    iterations after an untrained network."""
    ws = sc.WarmStart(11)
    ws.SetLearningRate(2e-3)
    for k, (r, y, young) in enumerate([(0.25, 0.4, 5e3), (0.35, 0.6, 2e4)]):
        sv = _small(n=48, sphere_radius=r, sdf_half=r + 0.1, offset=(0.0, y, 0.0), young=young)
        _, tp = S.build_patch(sv, sc)
        tp.GetModel().SetWarmStart(ws, "train")
        for _ in range(60):
            tp.Step(sv.dt)
    scene = _small(n=48)
    _, p = S.build_patch(scene, sc)
    m = p.GetModel()
    untrained = sc.WarmStart(12)
    m.SetWarmStart(ws, "track")  # the plain trajectory, corrections recorded
    res = {"plain4": [], "warm4": [], "untrained4": [], "plain20": []}
    dist = {"plain4": [], "warm4": [], "untrained4": []}
    for f in range(40):
        if f >= 30:
            m.save_state()
            x0 = m.positions().astype(np.float64)
            P = {}
            for key, setup in (
                    ("plain20", lambda mm: mm.SetWarmStart(ws, "track")),
                    ("plain4", lambda mm: (mm.SetWarmStart(ws, "track"), mm.SetNumIterations(4))),
                    ("warm4", lambda mm: (mm.SetNumIterations(scene.iterations), mm.SetWarmStart(ws, "solve", 4),
                                          mm.SetWarmStartFallback(1e9))),
                    ("untrained4", lambda mm: (mm.SetWarmStart(untrained, "solve", 4),
                                               mm.SetWarmStartFallback(1e9)))):
                out = []
                res[key].append(_residual_after(p, scene.dt, setup, out))
                P[key] = out[0]
            span = np.linalg.norm(P["plain20"] - x0)
            for key in dist:
                dist[key].append(np.linalg.norm(P[key] - P["plain20"]) / span)
            m.restore_state()
            m.SetWarmStart(ws, "track")
        p.Step(scene.dt)
    r = {k: float(np.mean(v)) for k, v in res.items()}
    d = {k: float(np.mean(v)) for k, v in dist.items()}
    print("mean residuals over frames 30-39:", r)
    print("mean distance to the plain 20-iteration frame, relative to the frame's motion:", d)
    assert r["plain20"] < r["plain4"]
    assert r["warm4"] < r["plain4"], r
    assert d["warm4"] < d["plain4"], d
    assert d["warm4"] < d["untrained4"], d
