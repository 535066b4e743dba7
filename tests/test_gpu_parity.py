"""GPU parity: the HIP substep through the C ABI against the CPU oracle run
with its constraints permuted into the GPU's Gauss-Seidel order.

Bar: per-vertex positions equal to the oracle's. The kernels and the oracle
evaluate the same FP32 operation sequence without contraction, and every
group of constraints that runs concurrently touches disjoint vertices, so
equality is expected to be bit-exact; the stated tolerance (1e-5 relative to
the cloth extent) is the acceptance bar and bit-exactness is asserted on top.
Every sweep schedule is covered: one launch per colour (mode 0), slab phases
staged in LDS (mode 1), with strips and interior/seam slabs of several sizes,
the triangles on tiles with halos (mode 3), whose vertex contacts of an
iteration without a bend pass are applied at the end of the FEM launch, and
mode 3's schedules packed for the kernels that request a launch's data before
its first colour (mode 4, the default).
"""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5
MODES = [0, 1, 3, 4]


def _pair(sc, oracle_mod, scene, mode=1):
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(mode)
    o = oracle_mod.from_scene(scene, tri_order=m.constraint_order("tri"), bend_order=m.constraint_order("bend"))
    return world, patch, m, o


def _compare(gpu, ora, what):
    ext = max(1e-6, float(np.abs(ora).max()))
    err = float(np.abs(gpu.astype(np.float64) - ora).max())
    assert np.isfinite(gpu).all(), what
    assert err <= REL_TOL * ext, f"{what}: max abs diff {err:.3e} > {REL_TOL} x {ext:.3f}"
    return err


@pytest.mark.parametrize("mode", MODES)
def test_drape32_100_substeps_bit_exact(sc, oracle_mod, mode):
    scene = S.DRAPE32
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(25):  # 25 frames x 4 substeps = 100 substeps
        patch.Step(scene.dt)
        o.step(scene.dt)
    gp = m.positions()
    op = o.get("pos")
    err = _compare(gp, op, "drape32 positions after 100 substeps")
    assert err == 0.0, f"expected bit-exact parity, max abs diff {err:.3e}"
    assert np.array_equal(m.velocities().astype(np.float64), o.get("vel"))


@pytest.mark.parametrize("mode,size", [(1, 1024), (1, 256), (2, 512), (2, 2048)])
def test_slab_shapes_bit_exact(sc, oracle_mod, mode, size):
    """Strips and interior/seam slabs of several sizes on a 96^2 drape (many
    slabs per phase, constraint records and vertices staged in LDS)."""
    scene = S.DRAPE32.with_(n=96)
    print(f"slab shapes mode={mode} size={size}", flush=True)
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(mode, size)
    o = oracle_mod.from_scene(scene, tri_order=m.constraint_order("tri"), bend_order=m.constraint_order("bend"))
    for f in range(6):
        patch.Step(scene.dt)
        o.step(scene.dt)
    err = _compare(m.positions(), o.get("pos"), f"96^2 drape, sweep {mode} slab size {size}")
    assert err == 0.0, f"expected bit-exact parity, max abs diff {err:.3e}"


@pytest.mark.parametrize("mode", MODES)
def test_drape32_reference_early_exit(sc, oracle_mod, mode):
    scene = S.DRAPE32.with_(fixed_iterations=False)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(10):
        patch.Step(scene.dt)
        o.step(scene.dt)
    assert m.stats()["iterations"] == o.stats()["iterations"]
    assert np.array_equal(m.positions().astype(np.float64), o.get("pos"))


@pytest.mark.parametrize("mode", MODES)
def test_drape_with_sdf_and_ground_reduced(sc, oracle_mod, mode):
    # the drape256 workload on a 48x48 grid, long enough to land on the
    # sphere and the ground
    scene = S.DRAPE256.with_(n=48)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(40):
        patch.Step(scene.dt)
        o.step(scene.dt)
    gp, op = m.positions(), o.get("pos")
    assert np.array_equal(gp.astype(np.float64), op), float(np.abs(gp - op).max())
    # contact sets of the last substep: identical as sets
    pts, nrm = m.contacts()
    ids, data = o.contacts()
    assert len(pts) == len(ids) and len(pts) > 0
    g = {tuple(np.r_[p, n].astype(np.float32).tolist()) for p, n in zip(pts, nrm)}
    r = {tuple(d[:6].astype(np.float32).tolist()) for d in data}
    assert g == r


@pytest.mark.parametrize("mode", MODES)
def test_drape_with_contacts_reference_early_exit(sc, oracle_mod, mode):
    """Early-exit mode keeps one contact launch per iteration (the exit test
    reads that iteration's contact error); the iteration counts and positions
    must follow the oracle while the cloth lies on the sphere and ground."""
    scene = S.DRAPE256.with_(n=40, fixed_iterations=False, iterations=12)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(45):
        patch.Step(scene.dt)
        o.step(scene.dt)
        assert m.stats()["iterations"] == o.stats()["iterations"], f
    assert np.array_equal(m.positions().astype(np.float64), o.get("pos"))
    assert len(o.contacts()[0]) > 0


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("iterations", [20, 7, 1])
def test_fixed_count_reports_last_pass_errors(sc, oracle_mod, mode, iterations):
    """With a fixed iteration count the reported bend and contact errors are
    those of the last bend pass and the last iteration's contacts, as the
    oracle's loop leaves them (ClothPBD.cpp:337-361), whichever launches
    carry them: equal to the bit while the cloth lies on the sphere. Catches
    a launch that skips its report in the last pass, or an earlier pass's
    error that survives into the report."""
    scene = S.DRAPE256.with_(n=40, iterations=iterations)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    seen = 0
    for f in range(40):
        patch.Step(scene.dt)
        o.step(scene.dt)
        gs, os_ = m.stats(), o.stats()
        assert np.float32(gs["bendError"]) == np.float32(os_["bendError"]), (f, gs, os_)
        assert np.float32(gs["contactError"]) == np.float32(os_["contactError"]), (f, gs, os_)
        seen += os_["contactError"] > 0
    assert seen > 0


@pytest.mark.parametrize("mode", MODES)
def test_drape256_full_size_parity(sc, oracle_mod, mode):
    scene = S.DRAPE256
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(3):
        patch.Step(scene.dt)
        o.step(scene.dt)
    gp, op = m.positions(), o.get("pos")
    _compare(gp, op, "drape256 after 12 substeps")
    assert np.array_equal(gp.astype(np.float64), op)


@pytest.mark.parametrize("mode", [1, 3, 4])
def test_drape256_slab_schedule_long_run(sc, oracle_mod, mode):
    """25 frames (100 substeps) of the full-size workload: the cloth reaches
    the sphere, so contacts and friction are active in the compared run."""
    scene = S.DRAPE256
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    for f in range(25):
        patch.Step(scene.dt)
        o.step(scene.dt)
    gp, op = m.positions(), o.get("pos")
    _compare(gp, op, "drape256 after 100 substeps")
    assert np.array_equal(gp.astype(np.float64), op)
    assert len(o.contacts()[0]) > 0


def test_graph_and_direct_launch_identical(sc):
    scene = S.DRAPE256.with_(n=40)
    out = []
    for use_graph in (True, False):
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        m.SetUseGraph(use_graph)
        for f in range(20):
            patch.Step(scene.dt)
        out.append(m.positions())
    assert np.array_equal(out[0], out[1])


def test_run_twice_deterministic(sc):
    scene = S.DRAPE256.with_(n=64)
    out = []
    for _ in range(2):
        world, patch = S.build_patch(scene, sc)
        for f in range(10):
            patch.Step(scene.dt)
        out.append(patch.GetModel().positions())
    assert np.array_equal(out[0], out[1])


def test_packed_sweeps_equal_unpacked(sc):
    """Mode 4 runs mode 3's schedules through the packed kernels: the same
    constraint order and the same bits, on the full-size drape while it lies
    on the sphere (contacts applied by FEM tiles and by bend slabs), with the
    solver statistics alike."""
    scene = S.DRAPE256
    out = []
    for mode in (3, 4):
        print(f"drape256 sweep {mode}", flush=True)
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        m.SetSweep(mode)
        for f in range(30):
            patch.Step(scene.dt)
        out.append((m.positions(), m.velocities(), m.constraint_order("tri"), m.constraint_order("bend"),
                    m.stats(), len(m.contacts()[0])))
    (p3, v3, t3, b3, s3, c3), (p4, v4, t4, b4, s4, c4) = out
    assert np.array_equal(t3, t4) and np.array_equal(b3, b4)
    assert c3 == c4 and c3 > 0
    assert np.array_equal(p3, p4)
    assert np.array_equal(v3, v4)
    assert s3 == s4


@pytest.mark.parametrize("flags", [0, S.CF_VERTICES | S.CF_EDGES])
def test_wide_packed_bend_slabs(sc, oracle_mod, flags):
    """A 400-vertex-wide cloth: its bend seams hold more than 1024 vertices,
    so mode 4 runs them on packed slabs of five or six vertex slots per
    thread, which apply no vertex contacts (without collision flags nothing
    to apply; with edge contacts the contact launches stay separate). Same
    bits as mode 3 and as the oracle."""
    scene = S.DRAPE32.with_(n=400, flags=flags)
    out = []
    for mode in (3, 4):
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        m.SetSweep(mode)
        for f in range(4):
            patch.Step(scene.dt)
        out.append((m.positions(), m.velocities(), m.constraint_order("bend")))
    (p3, v3, b3), (p4, v4, b4) = out
    assert np.array_equal(b3, b4)
    assert np.array_equal(p3, p4) and np.array_equal(v3, v4)
    o = oracle_mod.from_scene(scene, tri_order=m.constraint_order("tri"), bend_order=b4)
    for f in range(4):
        o.step(scene.dt)
    assert np.array_equal(p4.astype(np.float64), o.get("pos"))
