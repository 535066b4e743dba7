"""Host-side checks of libsolidcloth.so that need no GPU: the C ABI exports
every function include/solidcloth.h declares, cloth construction matches the
oracle bit for bit, the Gauss-Seidel colouring is a valid vertex-disjoint
partition, and misuse is reported instead of crashing. No compute call is made."""
import ctypes
import re
import subprocess

import numpy as np
import pytest

from solidcloth_amd import scenes as S


def header_functions(sc):
    text = open(sc.HEADER_PATH).read()
    return sorted(set(re.findall(r"SCEXP\s+[\w\s\*]+?\b(sc\w+)\s*\(", text)))


def test_library_exports_every_header_function(sc):
    declared = header_functions(sc)
    assert len(declared) >= 59 + 20
    out = subprocess.check_output(["nm", "-D", "--defined-only", sc.LIB_PATH]).decode()
    exported = set(re.findall(r"\bT (sc\w+)$", out, re.M))
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    L = sc.lib()
    for f in declared:
        assert getattr(L, f) is not None
    # the Python mirror binds exactly the declared functions
    assert sc.exported_names() == declared


def test_reference_abi_names_present(sc):
    """The 59 exports of the reference's ClothLib (Demo/ClothAPI.h:65-146)."""
    ref = """scCreateCollisionWorld scFreeCollisionWorld scClearCollidables scGetNumCollidables scAddCollidable
    scGetCollidable scCreateWalls scCreateSphere scCreateCapsule scCreateSDF scCreateMesh scGetCollidableType
    scGetMeshTree scGetWallsInfo scGetSphereInfo scGetCapsuleInfo scCreateClothPatch scFreeClothPatch
    scGetClothModel scGetClothMesh scInitCloth scStepCloth scUpdateClothMesh scComputeStrainMap
    scGetNumParticles scGetParticlePos scGetParticleInvMass scGetNumTris scGetTriangle scGetTriangleInfo
    scGetNumQuads scGetQuad scGetNumLinks scGetLink scGetNumContacts scGetContact scGetNumTriContacts
    scGetTriContact scGetCollisionFlags scSetCollisionFlags scGetThickness scSetThickness scGetTolerance
    scSetTolerance scGetFriction scSetFriction scGetNumIterations scSetNumIterations scGetUseFEM scSetUseFEM
    scGetClothTree scAddMouseSpring scUpdateMouseSpring scRemoveMouseSpring scGetMeshVertexCount
    scGetMeshIndexCount scGetMeshNormalBuffer scGetMeshVertexBuffer scGetMeshIndexBuffer""".split()
    assert len(ref) == 59
    assert set(ref) <= set(header_functions(sc))


@pytest.mark.parametrize("scene", [S.DRAPE32, S.DRAPE256.with_(n=40)], ids=lambda s: s.name)
def test_construction_matches_oracle(sc, oracle_mod, scene):
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    o = oracle_mod.from_scene(scene)
    assert np.array_equal(m.positions().astype(np.float64), o.get("pos"))
    assert np.array_equal(m.inv_masses().astype(np.float64), o.get("invMass"))
    assert np.array_equal(m.triangles(), o.triangles()[0])
    assert np.array_equal(m.bends(), o.bends())
    # rest state (ClothModel::AddTriangle) bit for bit: invDet, du, dv
    _, rest = o.triangles()
    L = sc.lib()
    inv = ctypes.c_float()
    du, dv, s = sc.scVector2(), sc.scVector2(), sc.scVector2()
    for k in range(0, len(rest), max(1, len(rest) // 97)):
        L.scGetTriangleInfo(m.h, k, ctypes.byref(inv), ctypes.byref(du), ctypes.byref(dv), ctypes.byref(s))
        assert inv.value == np.float32(rest[k, 5])
        assert (du.x, du.y, dv.x, dv.y) == tuple(np.float32(rest[k, [0, 1, 2, 3]]))


def test_reference_counts(sc):
    """Counts of the reference's grid construction at 32x32 and 256x256."""
    for scene, V, T, B, links in [(S.DRAPE32, 1024, 1922, 2821, 3906), (S.DRAPE256, 65536, 130050, 194565, None)]:
        world, patch = S.build_patch(scene, sc)
        m = patch.GetModel()
        assert (m.GetNumParticles(), m.GetNumTris(), m.GetNumBends()) == (V, T, B)
        if links:
            assert m.GetNumLinks() == links
        assert m.GetNumQuads() == (scene.n - 1) ** 2


def validate_schedule(order, offsets, sched, cons):
    """The race-freedom conditions of a parallel Gauss-Seidel schedule:
    every constraint exactly once; the order is sorted by (phase, slab,
    colour, index); slabs that run in the same phase touch disjoint vertex
    sets; constraints that run concurrently inside a slab (one colour) touch
    disjoint vertices. In colour mode (slab -1) a phase is one colour."""
    phase, slab, colour = sched
    assert sorted(order.tolist()) == list(range(len(cons)))
    keys = list(zip(phase.tolist(), slab.tolist(), colour.tolist(), order.tolist()))
    assert keys == sorted(keys)
    assert np.array_equal(np.searchsorted(phase, np.arange(len(offsets) - 1)), offsets[:-1])
    for p in range(len(offsets) - 1):
        sel = phase == p
        slabs = np.unique(slab[sel])
        owner = {}
        for s in slabs:
            in_slab = sel & (slab == s)
            for v in np.unique(cons[order[in_slab]]):
                assert owner.setdefault(int(v), int(s)) == s, f"phase {p}: slabs {owner[int(v)]} and {s} share {v}"
            for c in np.unique(colour[in_slab]):
                verts = cons[order[in_slab & (colour == c)]].ravel()
                assert len(np.unique(verts)) == len(verts), f"slab {s} colour {c} shares a vertex"


@pytest.mark.parametrize("n", [32, 97, 256])
@pytest.mark.parametrize("mode,size", [(0, 0), (1, 0), (1, 1024), (2, 1024), (2, 3072)])
def test_schedule_is_race_free(sc, n, mode, size):
    scene = S.DRAPE32.with_(n=n)
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(mode, size)
    for kind, cons in (("tri", m.triangles()), ("bend", m.bends())):
        validate_schedule(m.constraint_order(kind), m.colour_offsets(kind), m.schedule(kind), cons)
    # links (the non-FEM model's constraints) are scheduled the same way
    lm = S.build_patch(scene.with_(fem=False, dihedral=False), sc)[1].GetModel()
    lm.SetSweep(mode, size)
    validate_schedule(lm.constraint_order("link"), lm.colour_offsets("link"), lm.schedule("link"), lm.links()[0])
    nt = len(m.colour_offsets("tri")) - 1
    nb = len(m.colour_offsets("bend")) - 1
    if mode == 0:
        # a grid vertex of valence 8 makes 8 triangle colours the minimum; the
        # greedy colouring reaches it. Bends need at least 16.
        assert nt == 8 and 16 <= nb <= 20
    else:
        # a strip of slabs along the cloth needs two phases (one if it fits one
        # slab); thin bend slabs (a bend spans three grid rows) may need three;
        # interiors take one phase and seams one or two
        assert nt <= 2 and nb <= 3
        ph, sl, co = m.schedule("bend")
        assert co.max() + 1 <= 20


def test_seamed_schedule_shortens_the_sweep(sc):
    """On the 256^2 grid, interiors and seams need fewer sequential colour
    steps than strips (FEM 13 against 18, bends 23 against 35), and the
    library's choice is the one with the lower estimated cost."""
    world, patch = S.build_patch(S.DRAPE32.with_(n=256), sc)
    m = patch.GetModel()

    def steps(kind):
        ph, sl, co = m.schedule(kind)
        return sum(int(max(co[(ph == p) & (sl == s)].max() + 1 for s in np.unique(sl[ph == p])))
                   for p in np.unique(ph))

    m.SetSweep(1, 1024)
    strips = steps("tri"), steps("bend")
    m.SetSweep(2, 1024)
    seamed = steps("tri"), steps("bend")
    m.SetSweep(1, 0)
    auto = steps("tri"), steps("bend")
    assert seamed[0] < strips[0] and seamed[1] < strips[1], (strips, seamed)
    assert auto == seamed


def test_colouring_deterministic(sc):
    a = S.build_patch(S.DRAPE32, sc)[1].GetModel()
    b = S.build_patch(S.DRAPE32, sc)[1].GetModel()
    assert np.array_equal(a.constraint_order("bend"), b.constraint_order("bend"))


def test_errors_are_reported_not_fatal(sc):
    world = sc.CollisionWorld()
    patch = sc.ClothPatch(world)
    m = patch.GetModel()
    with pytest.raises(sc.SolidClothError):
        m.GetParticlePos(0)  # no cloth yet: index out of range
    with pytest.raises(sc.SolidClothError):
        patch.Init(1, 5, 0.1)
    with pytest.raises(sc.SolidClothError):
        patch.SetNumSteps(0)
    with pytest.raises(sc.SolidClothError):
        sc.SDF(np.zeros(4, np.float32), (0, 1, 1), (0, 0, 0), (1, 1, 1))  # zero cells
    L = sc.lib()
    n0 = L.scGetErrorCount()
    assert L.scGetNumParticles(sc.scModel()) == 0  # null handle
    assert L.scGetErrorCount() == n0 + 1
    assert b"null" in L.scGetLastError()


def test_collidable_introspection(sc):
    L = sc.lib()
    w = sc.CollisionWorld()
    hw = w.AddWalls((-1, -2, -3), (4, 5, 6))
    hs = w.AddSphere((0.5, 0.25, 0.0), 0.75)
    assert w.GetNumCollidables() == 2
    assert L.scGetCollidableType(hw) == sc.CT_WALLS and L.scGetCollidableType(hs) == sc.CT_SPHERE
    box = sc.scAabb3()
    L.scGetWallsInfo(hw, ctypes.byref(box))
    assert (box.min.x, box.min.y, box.min.z, box.max.x, box.max.y, box.max.z) == (-1, -2, -3, 4, 5, 6)
    r, c = ctypes.c_float(), sc.scVector3()
    L.scGetSphereInfo(hs, ctypes.byref(r), ctypes.byref(c))
    assert (r.value, c.x, c.y, c.z) == (0.75, 0.5, 0.25, 0.0)
    # type-mismatched getter leaves its outputs untouched (Demo/ClothAPI.cpp:92-124)
    r2 = ctypes.c_float(-1)
    L.scGetSphereInfo(hw, ctypes.byref(r2), ctypes.byref(c))
    assert r2.value == -1
    assert L.scGetCollidable(w.h, 1) == hs


def test_mesh_buffers_and_strain_map(sc):
    world, patch = S.build_patch(S.DRAPE32, sc)
    L = sc.lib()
    mesh = L.scGetClothMesh(patch.h)
    assert L.scGetMeshVertexCount(mesh) == 1024
    assert L.scGetMeshIndexCount(mesh) == 3 * 1922
    vb = L.scGetMeshVertexBuffer(mesh)
    v = np.ctypeslib.as_array(ctypes.cast(vb, ctypes.POINTER(ctypes.c_float)), shape=(1024 * 3,))
    assert np.array_equal(v.reshape(-1, 3), patch.GetModel().positions())
    nb = L.scGetMeshNormalBuffer(mesh)
    n = np.ctypeslib.as_array(ctypes.cast(nb, ctypes.POINTER(ctypes.c_float)), shape=(1024 * 3,)).reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    # rest state: zero strain everywhere
    colors = (sc.scVector3 * 1024)()
    L.scComputeStrainMap(patch.h, colors, 1024)
    assert max(abs(c.x) for c in colors) < 1e-4


@pytest.mark.parametrize("subdiv", [1, 3])
def test_obstacle_mesh_normals_and_edges_match_oracle(sc, oracle_mod, subdiv):
    """An obstacle mesh built through scCreateMeshFromArrays has the oracle's
    (reference-pinned) angle-weighted vertex normals, bit for bit."""
    v, t = S.icosphere(subdiv, 0.4, (0.1, -0.2, 0.05))
    mesh = sc.Mesh(v, t)
    L = sc.lib()
    nv = L.scGetMeshVertexCount(mesh.h)
    assert nv == len(v) and L.scGetMeshIndexCount(mesh.h) == t.size
    nb = L.scGetMeshNormalBuffer(mesh.h)
    n = np.ctypeslib.as_array(ctypes.cast(nb, ctypes.POINTER(ctypes.c_float)), shape=(nv * 3,)).reshape(-1, 3)
    assert np.array_equal(n, oracle_mod.mesh_normals(v, t))
    assert len(oracle_mod.construct_edges(t)) == 30 * 4 ** subdiv


@pytest.mark.parametrize("n", [16, 57])
def test_feature_contact_groups_are_vertex_disjoint(sc, n):
    """Obstacle-edge contacts are grouped by a colouring of the cloth edges,
    obstacle-vertex contacts by one of the cloth triangles; members of a group
    share no vertex, so they are solved in parallel in the order exported."""
    world, patch = S.build_patch(S.DRAPE32.with_(n=n), sc)
    m = patch.GetModel()
    edges = np.array(sorted({tuple(sorted((int(a), int(b)))) for tri in m.triangles()
                             for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0]))}))
    for kind, cons in (("edge_contact", edges), ("tri_contact", m.triangles())):
        order = m.constraint_order(kind)
        off = m.colour_offsets(kind)
        assert sorted(order) == list(range(len(cons)))
        assert off[0] == 0 and off[-1] == len(cons)
        for g in range(len(off) - 1):
            verts = cons[order[off[g]:off[g + 1]]].ravel()
            assert len(np.unique(verts)) == len(verts)
    assert len(m.colour_offsets("tri_contact")) - 1 == 8
    assert len(m.colour_offsets("edge_contact")) - 1 <= 11


def test_grid_sdf_edges_without_vertex_contacts_refused(sc):
    """The reference's grid-SDF segment query is a stub: without vertex
    contacts it yields undefined edge contacts, so the step refuses before
    touching the device."""
    scene = S.DRAPE256.with_(n=16, flags=S.CF_EDGES | S.CF_WALLS)
    world, patch = S.build_patch(scene, sc)
    with pytest.raises(sc.SolidClothError, match="grid SDF"):
        patch.Step(scene.dt)


def test_self_collision_parameters_validated(sc):
    world, patch = S.build_patch(S.FOLD24, sc)
    m = patch.GetModel()
    m.SetSelfParams(0.9, 0.003)
    m.SetSelfCapacity(100)
    with pytest.raises(sc.SolidClothError):
        m.SetSelfCCD(-1.0, 0.01, 0.1)
    with pytest.raises(sc.SolidClothError):
        m.SetSelfCapacity(-5)
    keys, ids, data = m.self_contacts("vt")  # before any step: none
    assert len(keys) == 0


def test_batch_scenes_follow_the_workload_description():
    scenes = S.batch_scenes(512)
    assert len(scenes) == 512
    ns = np.array([s.n for s in scenes])
    assert ns.min() >= 16 and ns.max() <= 128
    young = np.log10([s.young for s in scenes])
    assert young.min() >= 3 and young.max() <= 5
    assert all(0.1 <= s.poisson <= 0.45 for s in scenes)
    assert all(1e-3 <= s.bend <= 1e-1 for s in scenes)
    assert all(0.2 <= s.sphere_radius <= 0.5 and s.sdf_cells == 32 for s in scenes)
    for s in scenes:
        assert 1 <= len(s.pins) <= 4 and len(set(s.pins)) == len(s.pins)
        for v in s.pins:
            j, i = divmod(v, s.n)
            assert i in (0, s.n - 1) or j in (0, s.n - 1)
            for u in s.pins:
                if u != v:
                    assert max(abs(j - u // s.n), abs(i - u % s.n)) >= 2
    # seeded: the same list every time
    assert S.batch_scenes(8) == S.batch_scenes(8)


def test_batch_pins_and_validation(sc):
    """Pins become zero inverse masses; a batch refuses, before touching the
    device, patches it cannot step (self collision, mismatched substeps)."""
    scenes = S.batch_scenes(2, n_range=(16, 20))
    built = [S.build_patch(s, sc) for s in scenes]
    m = built[0][1].GetModel()
    im = m.inv_masses()
    assert set(np.flatnonzero(im == 0).tolist()) == set(scenes[0].pins)
    with pytest.raises(sc.SolidClothError):
        m.set_inv_masses(-np.ones(m.GetNumParticles()))
    b = sc.ClothBatch([p for _, p in built])
    assert b.GetNumPatches() == 2
    with pytest.raises(sc.SolidClothError):
        b.Add(built[0][1])  # already in the batch
    with pytest.raises(sc.SolidClothError):
        b.SetLaunch(300, 0)
    b.SetLaunch(512, 1)
    assert b.GetLaunch() == (512, 1, True)
    built[1][1].SetNumSteps(3)
    with pytest.raises(sc.SolidClothError, match="same number of substeps"):
        b.Step(1.0 / 60)
    built[1][1].SetNumSteps(4)
    world, patch = S.build_patch(S.FOLD24, sc)
    b.Add(patch)
    with pytest.raises(sc.SolidClothError, match="vertex and wall contacts"):
        b.Step(1.0 / 60)
    empty = sc.ClothPatch(sc.CollisionWorld())
    b2 = sc.ClothBatch([empty])
    with pytest.raises(sc.SolidClothError, match="no cloth"):
        b2.Step(1.0 / 60)


def _emulated_sweeps(tris, nv, ts, seed):
    """A deterministic nonlinear update of each triangle's three vertices,
    run (a) one triangle after another in the schedule's colour-major order
    and (b) tile by tile as the GPU runs it: each tile on a private copy of
    its vertices, all colours over its own instances, writing back only its
    own vertices. Returns both results."""
    rng = np.random.default_rng(seed)
    x0 = rng.normal(size=(nv, 2))

    def project(x, t):
        a, b, c = tris[t]
        m = (x[a] + x[b] + x[c]) / 3.0
        s = np.sin(x[a, 0] * 1.3 + x[b, 1] * 0.7 - x[c, 0] * 0.5) * 0.2
        x[a] = x[a] + s * (m - x[a]) + 0.01
        x[b] = x[b] + s * (m - x[b]) - 0.02
        x[c] = x[c] + 0.5 * s * (m - x[c])

    nc = ts["colours"]
    ntiles = len(ts["vert_begin"]) - 1
    seq = x0.copy()
    # colour-major order: every instance's triangle, each once, colour by colour
    col_of = {}
    for tile in range(ntiles):
        for col in range(nc):
            lo, hi = ts["col_offsets"][tile * nc + col], ts["col_offsets"][tile * nc + col + 1]
            for t in ts["instances"][lo:hi]:
                col_of[int(t)] = col
    order = sorted(col_of, key=lambda t: (col_of[t], t))
    for t in order:
        project(seq, t)
    tiled = x0.copy()
    for tile in range(ntiles):
        vb, ve = ts["vert_begin"][tile], ts["vert_begin"][tile + 1]
        verts, own = ts["verts"][vb:ve], ts["owned"][vb:ve]
        x = x0.copy()  # the tile stages its vertices from the state before the launch
        for col in range(nc):
            lo, hi = ts["col_offsets"][tile * nc + col], ts["col_offsets"][tile * nc + col + 1]
            for t in ts["instances"][lo:hi]:
                project(x, int(t))
        mine = verts[own == 1]
        tiled[mine] = x[mine]
    return seq, tiled, order


@pytest.mark.parametrize("n", [40, 97, 256])
def test_tile_schedule_equals_the_colour_major_sweep(sc, n):
    """Sweep mode 3 (FEM triangles on tiles with halos). Every vertex is one
    tile's own; each tile's instances of one colour share no vertex; a tile
    holds every triangle inside its extended region; and a nonlinear
    per-triangle update run tile by tile (private copies, own vertices
    written back) equals running it sequentially in the exported
    colour-major order, exactly, which holds only when every tile's halo
    covers its own vertices' dependencies over all colour steps."""
    world, patch = S.build_patch(S.DRAPE32.with_(n=n), sc)
    m = patch.GetModel()
    m.SetSweep(3, 0)
    ts = m.tile_schedule()
    assert ts is not None
    tris = m.triangles()
    nv = n * n
    owners = np.bincount(ts["verts"][ts["owned"] == 1], minlength=nv)
    assert np.all(owners == 1)
    nc = ts["colours"]
    assert nc == 8  # the grid's clique bound
    ntiles = len(ts["vert_begin"]) - 1
    for tile in range(ntiles):
        vb, ve = ts["vert_begin"][tile], ts["vert_begin"][tile + 1]
        region = set(ts["verts"][vb:ve].tolist())
        inst = []
        for col in range(nc):
            lo, hi = ts["col_offsets"][tile * nc + col], ts["col_offsets"][tile * nc + col + 1]
            vs = tris[ts["instances"][lo:hi]].ravel()
            assert len(vs) == len(set(vs.tolist())), (tile, col)
            inst.extend(ts["instances"][lo:hi].tolist())
        inside = [t for t in range(len(tris)) if set(tris[t].tolist()) <= region]
        assert sorted(inst) == inside, tile
    # the exported order is the colour-major order
    seq, tiled, order = _emulated_sweeps(tris, nv, ts, seed=n)
    assert list(m.constraint_order("tri")) == order
    assert np.array_equal(seq, tiled)


def test_model_getters_read_back_the_setters(sc):
    """The getters of Physics/ClothModel.h:110-146 that the library adds to
    the reference's ABI return what the setters stored (FP32)."""
    world, patch = S.build_patch(S.DRAPE32, sc)
    m = patch.GetModel()
    m.SetShearStiffness(123.5)
    m.SetStretchStiffness(0.75)
    m.SetBendStiffness(0.25)
    m.SetPoisson(0.5)
    m.SetDihedral(False)
    m.SetDragCoeff(0.875)
    m.SetSelfParams(0.625, 0.0078125)
    assert (m.GetShearStiffness(), m.GetStretchStiffness(), m.GetBendStiffness(), m.GetPoisson()) == \
        (123.5, 0.75, 0.25, 0.5)
    assert m.GetDihedral() is False and m.GetDragCoeff() == 0.875
    assert m.GetSelfParams() == (0.625, 0.0078125)


def test_bend_strips_keep_grid_rows_whole(sc):
    """The default bend schedule of the 256^2 grid: interiors of two-row
    strips (5 colours) and two seam phases (9 colours each, the most seam
    bends at one vertex), each strip made of whole rows, because strip cuts
    move back to the first vertex of a coordinate. Catches a cut inside a
    row (a strip holding part of a row, or more colours) and a cost model
    that picks the wider strips (15 unpaired interior colours)."""
    world, patch = S.build_patch(S.DRAPE256, sc)
    m = patch.GetModel()
    bends = m.bends()
    order = m.constraint_order("bend")
    ph, sl, co = m.schedule("bend")
    assert ph.max() == 2
    for p, want in ((0, 5), (1, 9), (2, 9)):
        assert co[ph == p].max() + 1 == want, p
    rows = bends[order] // 256
    for s in np.unique(sl[ph == 0]):
        r = rows[sl == s]
        # an interior slab: two whole rows 2k, 2k + 1
        assert r.min() % 2 == 0 and r.max() == r.min() + 1, s
