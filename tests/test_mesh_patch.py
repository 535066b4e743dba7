"""Mesh-based patches (ClothPatch::Init(const Mesh&), Physics/ClothPatch.cpp:
43-133) built by the library's host code and by the oracle: the same
welded particles, pins, triangles with their rest state, links, edges and
bends; the oracle's construction pinned by counts a direct enumeration
fixes — including the reference's missing closing edge of the last
triangle."""
import numpy as np
import pytest

from mesh_patch_util import seamed_grid


def _library_patch(sc, verts, tris, uvs, colors, attached=True):
    world = sc.CollisionWorld()
    patch = sc.ClothPatch(world)
    mesh = sc.Mesh(verts, tris)
    mesh.set_uvs(uvs)
    mesh.set_colors(colors)
    patch.InitFromMesh(mesh, (0.0, 0.3, 0.0), attached)
    return world, patch, patch.GetModel()


def test_mesh_patch_matches_oracle(sc, oracle_mod):
    verts, tris, uvs, colors = seamed_grid(9)
    world, patch, m = _library_patch(sc, verts, tris, uvs, colors)
    o = oracle_mod.Cloth()
    mp = o.init_mesh(verts, tris, uvs, colors, (0.0, 0.3, 0.0), True)
    assert m.GetNumParticles() == o.counts()["V"] == 81  # the seam's 9 copies welded
    assert np.array_equal(m.positions().astype(np.float64), o.get("pos"))
    assert np.array_equal(m.inv_masses().astype(np.float64), o.get("invMass"))
    assert (m.inv_masses() == 0).sum() == 2
    oi, orr = o.triangles()
    assert np.array_equal(m.triangles(), oi)
    assert np.array_equal(m.bends(), o.bends())
    li, lr = m.links()
    oli, olr = o.links()
    assert np.array_equal(li, oli) and np.allclose(lr, olr)


def test_mesh_patch_topology_counts(oracle_mod):
    verts, tris, uvs, colors = seamed_grid(7)
    o = oracle_mod.Cloth()
    mp = o.init_mesh(verts, tris, uvs, colors, (0, 0, 0), False)
    welded = mp[tris]
    # unique undirected edges of the welded triangles
    edges = set()
    for t in welded:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            edges.add((min(a, b), max(a, b)))
    last = welded[-1]
    closing = (min(last[2], last[0]), max(last[2], last[0]))
    occurrences = sum(1 for t in welded for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))
                      if (min(a, b), max(a, b)) == closing)
    expect_edges = len(edges) - (1 if occurrences == 1 else 0)
    c = o.counts()
    assert c["T"] == len(tris)
    assert c["E"] == expect_edges
    assert c["L"] == expect_edges  # one stretch link per edge (no pinned pair here)
    interior = sum(1 for e in edges
                   if sum(1 for t in welded for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))
                          if (min(a, b), max(a, b)) == e) == 2)
    assert c["B"] == interior - (1 if occurrences == 2 else 0) or c["B"] == interior
