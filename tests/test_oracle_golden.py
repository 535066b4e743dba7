"""The oracle against golden vectors written by the reference's own Geometry
code (tests/golden/make_golden.py compiles Geometry/Mesh.cpp, Collision3D.cpp,
AabbTree.cpp, ClosestPointOnMesh.cpp and VoronoiRegions.cpp from the
reference sources)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
MESHES = np.load(os.path.join(GOLDEN, "ref_meshes.npz"))


@pytest.mark.parametrize("name", ["grid32", "grid16", "ico1"])
def test_construct_edges_matches_reference(oracle_mod, name):
    """Mesh::ConstructEdges (Geometry/Mesh.cpp:65-147): unique sorted edges and
    their winged triangles t1/t2, including which triangle std::sort leaves
    first among duplicates."""
    got = oracle_mod.construct_edges(MESHES[f"{name}_indices"])
    exp = MESHES[f"{name}_edges"]
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


def test_grid_topology_matches_reference_mesh(oracle_mod):
    """ClothPatch::Init's grid (Physics/ClothPatch.cpp:135-246) at 32x32 yields
    the reference's triangle list and the edges the reference builds from it."""
    c = oracle_mod.Cloth().init_grid(32, 32, np.float32(1 / 31), (0, 0, 0), True, False)
    tris, _ = c.triangles()
    assert np.array_equal(tris.astype(np.uint32), MESHES["grid32_indices"])
    assert np.array_equal(c.edges(), MESHES["grid32_edges"])
    pos = c.get("pos").astype(np.float32)
    assert np.array_equal(pos, MESHES["grid32_vertices"])


COLL = np.load(os.path.join(HERE, "golden", "ref_collision3d.npz"))


def test_closest_pt_point_triangle_matches_reference(oracle_mod):
    """ClosestPtPointTriangle (Geometry/Collision3D.cpp:62-141): region, point
    and barycentric coordinates bit for bit, including vertex, edge and
    degenerate cases."""
    P = COLL["points"]
    for k in range(len(P)):
        a, b, c, q = P[k]
        r, cp, bar = oracle_mod.closest_pt_point_triangle(q, a, b, c)
        assert r == COLL["tri_region"][k], k
        assert np.array_equal(np.r_[cp, bar], COLL["tri_out"][k]), k


def test_closest_pt_segm_segm_matches_reference(oracle_mod):
    """ClosestPtSegmSegm (Geometry/Collision3D.cpp:10-58), bit for bit."""
    P = COLL["points"]
    for k in range(len(P)):
        a, b, c, q = P[k]
        r, s, t, c1, c2 = oracle_mod.closest_pt_segm_segm(a, b, c, q)
        assert r == COLL["segm_region"][k], k
        got = np.r_[s, t, c1, c2].astype(np.float32)
        exp = COLL["segm_out"][k]
        assert np.array_equal(got, exp) or (np.isnan(got).sum() == np.isnan(exp).sum()
                                             and np.array_equal(got[~np.isnan(got)], exp[~np.isnan(exp)])), k


def test_signed_volume_matches_reference(oracle_mod):
    P = COLL["points"]
    for k in range(len(P)):
        assert oracle_mod.signed_volume(*P[k]) == COLL["signed_volume"][k], k


@pytest.mark.parametrize("name", ["grid16", "ico1"])
def test_mesh_normals_match_reference(oracle_mod, name):
    """Angle-weighted vertex normals (Geometry/Mesh.cpp:10-63) bit for bit
    with the same libm acosf."""
    got = oracle_mod.mesh_normals(MESHES[f"{name}_vertices"], MESHES[f"{name}_indices"])
    assert np.array_equal(got, MESHES[f"{name}_normals"])


def test_bends_from_reference_edges(oracle_mod):
    """Each interior edge (t2 >= 0) of the reference's edge list gives one bend
    (i1, i2, opposite vertex of t1, opposite vertex of t2),
    Physics/ClothModel.cpp:94-126."""
    c = oracle_mod.Cloth().init_grid(32, 32, np.float32(1 / 31), (0, 0, 0), True, False)
    tris = MESHES["grid32_indices"]
    edges = MESHES["grid32_edges"]
    exp = []
    for i1, i2, t1, t2 in edges:
        if t2 < 0:
            continue
        o1 = [v for v in tris[t1] if v not in (i1, i2)][0]
        o2 = [v for v in tris[t2] if v not in (i1, i2)][0]
        exp.append((i1, i2, o1, o2))
    assert np.array_equal(c.bends(), np.array(exp, dtype=np.int32))
    assert len(exp) == 2821


def test_mesh_queries_match_the_compiled_reference(oracle_mod):
    """Vertex -> nearest obstacle triangle (ClosestPointOnMeshToPointAcc over
    the reference's own AabbTree, Geometry/ClosestPointOnMesh.cpp:304-319 and
    AabbTree.cpp), the solver's interpolated triangle normal
    (ComputeTriangleNormal, :37-62, as ClothCollisionHandler.cpp:159 calls it)
    and the edge normal (ComputeEdgeNormal, :483-497), against the compiled
    reference (tests/golden/ref_mesh_queries.npz; tie-free queries only). The
    oracle answers the vertex queries through its contact detection (query
    from prev, normal = delta / distance): triangle, closest point and normal
    must be the reference's bit for bit."""
    g = np.load(os.path.join(GOLDEN, "ref_mesh_queries.npz"))
    v, t, Q, ref = g["vertices"], g["indices"], g["queries"], g["query_out"]
    n = len(Q)
    pos = np.vstack([Q.astype(np.float64), [[3, 3, 3], [3.1, 3, 3], [3, 3, 3.1]]])
    c = oracle_mod.Cloth()
    c.set(thickness=1e-4, tolerance=1.0, collFlags=1, fem=1)
    c.init_arrays(pos, pos[:, [0, 2]], np.ones(len(pos)), [[n, n + 1, n + 2]])
    c.add_mesh(v, t)
    c.put("prev", pos)
    c.detect()
    ids, data = c.contacts()
    feats = c.contact_features()
    got = {int(i): (int(f), d) for i, f, d in zip(ids[:, 0], feats[:, 1], data)}
    assert sorted(got) == list(range(n))
    for k in range(n):
        tri, d = got[k]
        assert tri == int(ref[k, 0]), k
        assert np.array_equal(d[0:3].astype(np.float32), ref[k, 1:4]), k
        assert np.array_equal(d[3:6].astype(np.float32), ref[k, 4:7]), k
    tn, en, el = c.mesh_normals_at(0, g["tri"], g["coords"], g["edges"], g["params"])
    assert np.array_equal(tn, g["tri_normals"])
    assert np.array_equal(en, g["edge_normals"])
    # parameters exactly 0 and 1 took the ends' vertex normals
    assert not np.array_equal(en[0::3], en[2::3])


LIVE = np.load(os.path.join(GOLDEN, "ref_live_self.npz"))


@pytest.mark.parametrize("case", range(int(LIVE["cases"])))
def test_live_self_path_matches_the_compiled_reference(oracle_mod, case):
    """The brute-force live self path (MeshClosestPoints::ClosestPoints on the
    cloth's own mesh, Geometry/ClosestPointsBetweenMeshes.cpp:538-592) against
    the compiled reference, bit for bit and in order, on folded sheets, a
    squashed irregular mesh, larger copies of both and two crossing sheets
    (tests/golden/make_golden.py). Each of these mistakes, planted in the
    restatement, fails at least one case: the edge's orientation flag
    ignored, edge-region eps 0 instead of 0.002, vertex-region eps 0.01
    instead of 0.1, the border-vertex rule dropped, the duplicate-feature
    check dropped, an edge adjacency rule dropped, the vertex-region test run
    at the vertex instead of at the closest point, the EE search over all
    edges instead of the later ones, a same-triangle EE rejection dropped.
    Two forms are equivalent and so not told apart: the face test's chained
    (signbit == signbit) == signbit and "all three signs equal" (a point
    whose closest point is a face interior has all three signs negative), and
    the border-edge rule (a vertex nearest a border edge's interior lies on
    its outer side, where the edge-region test already passes)."""
    g = lambda k: LIVE[f"c{case}_{k}"]
    live = oracle_mod.LiveSelf(g("vertices"), g("indices"))
    vi, vd = live.vertex_queries(float(g("rad")))
    ei, ed = live.edge_queries()
    assert len(g("vt_ids")) > 0 and len(g("ee_ids")) > 0
    assert np.array_equal(vi, g("vt_ids"))
    assert np.array_equal(vd.view(np.uint32), g("vt_data").view(np.uint32))
    assert np.array_equal(ei, g("ee_ids"))
    assert np.array_equal(ed.view(np.uint32), g("ee_data").view(np.uint32))
    # region types: faces, edges and vertices all occur across the cases
    if case == 1:
        assert set(np.unique(vi[:, 3])) == {0, 1, 2}


def test_live_self_threads_and_subsets(oracle_mod):
    """Any thread count gives the serial order, and a query list gives the
    infos of those items only, in list order."""
    g = lambda k: LIVE[f"c1_{k}"]
    live = oracle_mod.LiveSelf(g("vertices"), g("indices"))
    try:
        oracle_mod.set_threads(4)
        vi, vd = live.vertex_queries(float(g("rad")))
        ei, ed = live.edge_queries()
    finally:
        oracle_mod.set_threads(1)
    assert np.array_equal(vi, g("vt_ids")) and np.array_equal(ei, g("ee_ids"))
    q = np.array([100, 3, 57], np.int32)
    si, _ = live.vertex_queries(float(g("rad")), q)
    want = np.concatenate([g("vt_ids")[g("vt_ids")[:, 0] == k] for k in q])
    assert np.array_equal(si, want)
    qe = np.array([5, 200, 11], np.int32)
    se, _ = live.edge_queries(qe)
    want = np.concatenate([g("ee_ids")[g("ee_ids")[:, 0] == k] for k in qe])
    assert np.array_equal(se, want)


SDFM = np.load(os.path.join(GOLDEN, "ref_sdf_from_mesh.npz"))


@pytest.mark.parametrize("case", range(int(SDFM["cases"])))
def test_sdf_from_mesh_matches_the_compiled_reference(oracle_mod, case):
    """SDF::CreateFromMesh (Geometry/SDF.cpp:58-76) restated against the
    reference's own tree and closest-point query at every node, bit for bit.
    The reference's query returns the unit delta as its normal, so its sign
    test never flips a value: every value is >= 0 even inside the closed
    meshes (a restatement that signed the field with a triangle or vertex
    normal fails cases 0 and 1). Case 0's unequal steps tell
    (max - min) * (1 / n) from (max - min) / n; case 2's nodes on the flat mesh
    skip the triangles at distance 0 (a restatement that kept them would give
    0 there)."""
    g = lambda k: SDFM[f"c{case}_{k}"]
    box = g("box")
    got = oracle_mod.sdf_from_mesh(g("vertices"), g("indices"), g("cells"), box[:3], box[3:])
    want = g("values")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (want >= 0).all()
    if case == 2:
        # nodes on the mesh plane (y = 0): the nearest triangle not through the node
        assert want[:, 2, :].min() > 0.3
